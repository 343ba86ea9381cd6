/* Double-precision 3-vector helpers used by the public headers and by callers.
 *
 * Same names and semantics as the reference's src/vec3.h:9-58 (callers such as the
 * driver and the reference tests use them directly), so they are kept alike:
 * v3_normalize scales by the reciprocal length; sphere_move is the Rodrigues-form
 * great-circle step that returns p unchanged for a vanishing tangent velocity.
 */
#ifndef GGB_VEC3_H
#define GGB_VEC3_H

#include <math.h>

typedef struct {
	double x, y, z;
} vec3;

static inline vec3 v3(double x, double y, double z)
{
	vec3 r;
	r.x = x;
	r.y = y;
	r.z = z;
	return r;
}
static inline vec3 v3_add(vec3 a, vec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline vec3 v3_sub(vec3 a, vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline vec3 v3_scale(vec3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline double v3_dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline vec3 v3_cross(vec3 a, vec3 b)
{
	return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double v3_len(vec3 a) { return sqrt(v3_dot(a, a)); }
static inline vec3 v3_normalize(vec3 a)
{
	double l = v3_len(a);
	if (l > 0.0)
		return v3_scale(a, 1.0 / l);
	return a;
}
/* tangent-plane projection: v minus its component along the unit normal n */
static inline vec3 v3_tangent(vec3 v, vec3 n) { return v3_sub(v, v3_scale(n, v3_dot(v, n))); }
/* angle between a and b, accurate for small angles */
static inline double v3_angle(vec3 a, vec3 b) { return atan2(v3_len(v3_cross(a, b)), v3_dot(a, b)); }
static inline vec3 v3_load(const float *f) { return v3(f[0], f[1], f[2]); }
static inline void v3_store(float *f, vec3 a)
{
	f[0] = (float) a.x;
	f[1] = (float) a.y;
	f[2] = (float) a.z;
}

/* Move unit vector p along the great circle of tangent velocity v for time s. */
static inline vec3 sphere_move(vec3 p, vec3 v, double s)
{
	vec3 t = v3_tangent(v, p);
	double speed = v3_len(t);
	if (speed < 1e-300)
		return p;
	double ang = speed * s;
	return v3_normalize(v3_add(v3_scale(p, cos(ang)), v3_scale(t, sin(ang) / speed)));
}

/* +Y is north, longitude 0 at +Z, increasing toward +X (reference vec3.h:44-58). */
static inline void dir_to_lonlat(vec3 d, double *lon, double *lat)
{
	double y = d.y;
	*lon = atan2(d.x, d.z);
	if (y > 1.0)
		y = 1.0;
	if (y < -1.0)
		y = -1.0;
	*lat = asin(y);
}

static inline vec3 lonlat_to_dir(double lon, double lat)
{
	return v3(cos(lat) * sin(lon), sin(lat), cos(lat) * cos(lon));
}

#endif
