/* Equiangular cubed-sphere grid: geometry, Voronoi stencil tables and sampling.
 *
 * Drop-in for the reference's src/cubesphere.h:29-73.  The struct keeps every field
 * callers read (N, ncells, h, pos, area, nnb, nbr, w, elen, gc, total_area,
 * asym_count, open_cells; read by main.c:390-392 and tests/test_cubesphere.c:48-75)
 * with the same meaning and layout conventions: cell c = (f*N + j)*N + i, neighbour
 * tables ELL-padded to CS_MAXNB per cell, float storage.  The trailing `dev` field
 * caches the device copy of the tables; it is managed by the library.
 *
 * Face basis (cubesphere.c:8-10): 0 = +Z, 1 = +X, 2 = -Z, 3 = -X, 4 = +Y, 5 = -Y.
 */
#ifndef GGB_CUBESPHERE_H
#define GGB_CUBESPHERE_H

#include "vec3.h"

#ifdef __cplusplus
extern "C" {
#endif

extern const vec3 cs_face_center[6];
extern const vec3 cs_face_right[6];
extern const vec3 cs_face_up[6];

#define CS_MAXNB 8

struct ggb_grid_dev;

typedef struct {
	int N;              /* cells per face edge */
	int ncells;         /* 6*N*N */
	double h;           /* angular cell width (pi/2)/N */
	float *pos;         /* [ncells*3] unit cell centres */
	float *area;        /* [ncells] Voronoi cell areas */
	unsigned char *nnb; /* [ncells] Voronoi neighbour counts */
	int *nbr;           /* [ncells*CS_MAXNB] neighbour indices, longest edge first */
	float *w;           /* [ncells*CS_MAXNB] Laplacian weights edge_len / centre_dist */
	float *elen;        /* [ncells*CS_MAXNB] Voronoi edge lengths (radians) */
	float *gc;          /* [ncells*CS_MAXNB*3] least-squares gradient coefficients */
	double total_area;
	int asym_count;     /* non-reciprocal Voronoi links (0 expected) */
	int open_cells;     /* Voronoi cells not closed by the candidates (0 expected) */
	int *ghost;         /* [6*(N+2)*(N+2)] tap index of every (face, i, j), i, j in [-1, N] */
	struct ggb_grid_dev *dev; /* device copy (library-managed) */
} CubeSphere;

/* reference cubesphere.c:233-359; returns 0, or -1 on allocation failure */
int cs_init(CubeSphere *cs, int N);
/* reference cubesphere.c:361-371 */
void cs_free(CubeSphere *cs);

/* reference cubesphere.c:12-20 */
int cs_face_of(vec3 d);
/* reference cubesphere.c:22-27 */
vec3 cs_face_ab_to_dir(int f, double a, double b);
/* reference cubesphere.c:29-38: continuous cell coordinates, centres at integers */
void cs_locate(const CubeSphere *cs, vec3 d, int *f, double *x, double *y);
/* reference cubesphere.c:40-51 */
int cs_cell_of(const CubeSphere *cs, vec3 d);
/* reference cubesphere.c:68-92: 4 bilinear taps and weights at direction p */
void cs_taps(const CubeSphere *cs, vec3 p, int idx[4], double wt[4]);
/* reference cubesphere.c:94-100 */
double cs_sample_scalar(const CubeSphere *cs, const float *fld, vec3 p);
/* reference cubesphere.c:102-111: bilinear vector sample projected to the tangent at p */
vec3 cs_sample_vec(const CubeSphere *cs, const float *fld3, vec3 p);

static inline int cs_index(const CubeSphere *cs, int f, int i, int j)
{
	return (f * cs->N + j) * cs->N + i;
}

static inline vec3 cs_pos(const CubeSphere *cs, int c) { return v3_load(cs->pos + 3 * c); }

#ifdef __cplusplus
}
#endif

#endif
