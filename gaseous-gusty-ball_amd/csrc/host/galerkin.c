/*
 * galerkin.c -- the coarse operators of the mixed-precision multigrid hierarchy.
 *
 * The reference coarsens by rediscretisation: every level is the Voronoi Laplacian of
 * its own N/2^l grid (fluid.c:41-98).  The mixed-precision preconditioner, which has no
 * counterpart in the reference, uses Galerkin operators instead, K_(l+1) = R_l K_l P_l with
 * the level's bilinear prolongation P_l and R_l = P_l^T: the coarse correction is then the
 * energy-norm projection, and with the same smoothing the preconditioned CG needs about a
 * third fewer iterations (the float64 restatement of the cycle measures a residual
 * reduction of ~100x per iteration against ~16x with the rediscretised levels).
 *
 * On the face interior the product is a 5x5 stencil.  Each operator is kept in the same
 * form as the fine one, (K x)_i = sum_k w_ik (x_i - x_j(k)) with w_ik = -(R K P)_ij, so
 * its row sums are zero by construction and the constant stays its null space; the
 * weights are symmetrised (w_ij and w_ji from the same two values) and rounded to half
 * floats, which the device stores.  The next level's product is formed from the rounded
 * weights, so every level is the Galerkin product of the operator the device applies.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ggb_host.h"

/* IEEE binary16 bits of a double, rounded to nearest even directly from the double */
unsigned short ggb_f16_bits(double v)
{
	uint64_t u;
	memcpy(&u, &v, 8);
	unsigned short sign = (unsigned short) ((u >> 48) & 0x8000u);
	int e = (int) ((u >> 52) & 0x7ff);
	uint64_t m = u & 0xfffffffffffffULL;
	if (e == 0x7ff)
		return (unsigned short) (sign | 0x7c00u | (m ? 0x200u : 0u));
	if (e == 0)
		return sign; /* double subnormals are far below half precision */
	m |= 1ULL << 52;
	int he = e - 1023 + 15;
	if (he >= 31)
		return (unsigned short) (sign | 0x7c00u);
	/* q: the value in units of the half's last place (2^(he - 25), or 2^-24 below the
	 * normal range); a carry out of the mantissa moves to the next binade by itself */
	int shift = he >= 1 ? 42 : 42 + 1 - he;
	if (shift > 54)
		return sign;
	uint64_t q = m >> shift, rem = m & ((1ULL << shift) - 1), half = 1ULL << (shift - 1);
	if (rem > half || (rem == half && (q & 1)))
		q++;
	if (he >= 1) {
		/* q holds the implicit bit: 0x400 .. 0x800 */
		if (q == 0x800u) {
			q = 0x400u;
			he++;
			if (he >= 31)
				return (unsigned short) (sign | 0x7c00u);
		}
		return (unsigned short) (sign | ((unsigned) he << 10) | (unsigned) (q & 0x3ffu));
	}
	return (unsigned short) (sign | (unsigned) q); /* q == 0x400: the smallest normal */
}

double ggb_f16_value(unsigned short h)
{
	int e = (h >> 10) & 0x1f, m = h & 0x3ff;
	double v;
	if (e == 0)
		v = ldexp((double) m, -24);
	else if (e == 31)
		v = m ? NAN : INFINITY;
	else
		v = ldexp((double) (m | 0x400), e - 25);
	return (h & 0x8000) ? -v : v;
}

void ggb_galop_free(GalOp *g)
{
	free(g->nbr);
	free(g->w);
	free(g->wh);
	free(g->nn);
	free(g->reg);
	free(g->dinv);
	memset(g, 0, sizeof(*g));
}

/* the fixed slot of offset (dy, dx), |dy|, |dx| <= 2, not both zero: the 5x5 square
 * row-major with the centre skipped */
static int wide_slot(int dy, int dx)
{
	int q = (dy + 2) * 5 + (dx + 2);
	return q < 12 ? q : q - 1;
}

#define ROWCAP 128  /* couplings of a row while it is summed */
#define ROWSTORE 48 /* off-diagonal couplings kept per row */

typedef struct {
	int cnt;
	int key[ROWCAP];
	double val[ROWCAP];
} Acc;

static void acc_add(Acc *a, int k, double v)
{
	for (int e = 0; e < a->cnt; e++)
		if (a->key[e] == k) {
			a->val[e] += v;
			return;
		}
	if (a->cnt == ROWCAP)
		ggb_die("fluid: a Galerkin row exceeds %d entries", ROWCAP);
	a->key[a->cnt] = k;
	a->val[a->cnt] = v;
	a->cnt++;
}

int ggb_galerkin(const CubeSphere *cs, const GalOp *gf, const int *ptap, const float *pw, const int *rstart,
		 const int *ridx, const float *rw, int Nc, GalOp *out)
{
	int nc = 6 * Nc * Nc;
	/* unsymmetrised rows: off-diagonal (J, -K_IJ) pairs of every coarse row */
	int *rcnt = calloc((size_t) nc, sizeof(int));
	int *rkey = malloc(sizeof(int) * (size_t) nc * ROWSTORE);
	double *rval = malloc(sizeof(double) * (size_t) nc * ROWSTORE);
	if (!rcnt || !rkey || !rval)
		ggb_die("fluid: out of memory");
	#pragma omp parallel
	{
		Acc *a = malloc(sizeof(Acc));
		if (!a)
			ggb_die("fluid: out of memory");
		#pragma omp for schedule(dynamic, 256)
		for (int I = 0; I < nc; I++) {
			a->cnt = 0;
			for (int e = rstart[I]; e < rstart[I + 1]; e++) {
				int i = ridx[e];
				double r = rw[e];
				int nn = gf ? gf->nn[i] : cs->nnb[i];
				const double *gw = gf ? gf->w + (size_t) GGBK_WNB * i : NULL;
				const int *gn = gf ? gf->nbr + (size_t) GGBK_WNB * i : cs->nbr + (size_t) CS_MAXNB * i;
				/* r (K e_J)_i summed over the taps: the diagonal sum_k w_ik times P(i, .),
				 * minus w_ik P(j, .) for each neighbour */
				double sw = 0.0;
				for (int k = 0; k < nn; k++)
					sw += gf ? gw[k] : (double) cs->w[(size_t) CS_MAXNB * i + k];
				for (int t = 0; t < 4; t++)
					acc_add(a, ptap[4 * (size_t) i + t], r * sw * pw[4 * (size_t) i + t]);
				for (int k = 0; k < nn; k++) {
					double wk = gf ? gw[k] : (double) cs->w[(size_t) CS_MAXNB * i + k];
					if (wk == 0.0)
						continue;
					int j = gn[k];
					double c = r * wk;
					for (int t = 0; t < 4; t++)
						acc_add(a, ptap[4 * (size_t) j + t], -c * pw[4 * (size_t) j + t]);
				}
			}
			int m = 0;
			for (int e = 0; e < a->cnt; e++) {
				if (a->key[e] == I)
					continue;
				if (m == ROWSTORE) {
					m = -1;
					break;
				}
				rkey[(size_t) I * ROWSTORE + m] = a->key[e];
				rval[(size_t) I * ROWSTORE + m] = -a->val[e];
				m++;
			}
			rcnt[I] = m;
		}
		free(a);
	}
	memset(out, 0, sizeof(*out));
	out->n = nc;
	out->N = Nc;
	out->nbr = malloc(sizeof(int) * (size_t) nc * GGBK_WNB);
	out->w = calloc((size_t) nc * GGBK_WNB, sizeof(double));
	out->wh = calloc((size_t) nc * GGBK_WNB, sizeof(unsigned short));
	out->nn = calloc((size_t) nc, 1);
	out->reg = calloc((size_t) nc, 1);
	out->dinv = malloc(sizeof(float) * (size_t) nc);
	if (!out->nbr || !out->w || !out->wh || !out->nn || !out->reg || !out->dinv)
		ggb_die("fluid: out of memory");
	int rc = 0;
	#pragma omp parallel for schedule(dynamic, 256) reduction(| : rc)
	for (int I = 0; I < nc; I++) {
		int f = I / (Nc * Nc), y = (I / Nc) % Nc, x = I % Nc;
		int *nb = out->nbr + (size_t) GGBK_WNB * I;
		double *w = out->w + (size_t) GGBK_WNB * I;
		unsigned short *wh = out->wh + (size_t) GGBK_WNB * I;
		/* symmetrised and rounded weights of the row's nonzero couplings */
		int cnt = 0, ok = 1;
		int jj[ROWSTORE];
		unsigned short hh[ROWSTORE];
		if (rcnt[I] < 0) {
			rc |= 4;
			continue;
		}
		for (int e = 0; e < rcnt[I]; e++) {
			int J = rkey[(size_t) I * ROWSTORE + e];
			double wij = rval[(size_t) I * ROWSTORE + e], wji = 0.0;
			int found = 0;
			for (int q = 0; q < rcnt[J]; q++)
				if (rkey[(size_t) J * ROWSTORE + q] == I) {
					wji = rval[(size_t) J * ROWSTORE + q];
					found = 1;
					break;
				}
			if (!found) {
				ok = 0;
				break;
			}
			unsigned short h = ggb_f16_bits(0.5 * (wij + wji));
			if ((h & 0x7fff) == 0)
				continue;
			jj[cnt] = J;
			hh[cnt] = h;
			cnt++;
		}
		if (!ok) {
			rc |= 1;
			continue;
		}
		/* regular: every coupling on the cell's face within two cells each way, and the
		 * whole 5x5 square on the face */
		int reg = x >= 2 && x <= Nc - 3 && y >= 2 && y <= Nc - 3;
		for (int e = 0; reg && e < cnt; e++) {
			int J = jj[e], fJ = J / (Nc * Nc), yJ = (J / Nc) % Nc, xJ = J % Nc;
			if (fJ != f || abs(yJ - y) > 2 || abs(xJ - x) > 2)
				reg = 0;
		}
		for (int k = 0; k < GGBK_WNB; k++)
			nb[k] = I;
		double sw = 0.0;
		if (reg) {
			for (int dy = -2; dy <= 2; dy++)
				for (int dx = -2; dx <= 2; dx++)
					if (dy || dx)
						nb[wide_slot(dy, dx)] = (f * Nc + y + dy) * Nc + x + dx;
			for (int e = 0; e < cnt; e++) {
				int J = jj[e], s = wide_slot((J / Nc) % Nc - y, J % Nc - x);
				wh[s] = hh[e];
			}
			out->nn[I] = GGBK_WNB;
		} else {
			if (cnt > GGBK_WNB) {
				rc |= 2;
				continue;
			}
			/* ascending neighbour index */
			for (int a = 1; a < cnt; a++)
				for (int b = a; b > 0 && jj[b - 1] > jj[b]; b--) {
					int tj = jj[b];
					jj[b] = jj[b - 1];
					jj[b - 1] = tj;
					unsigned short th = hh[b];
					hh[b] = hh[b - 1];
					hh[b - 1] = th;
				}
			for (int e = 0; e < cnt; e++) {
				nb[e] = jj[e];
				wh[e] = hh[e];
			}
			out->nn[I] = (unsigned char) cnt;
		}
		out->reg[I] = (unsigned char) reg;
		for (int k = 0; k < GGBK_WNB; k++) {
			w[k] = ggb_f16_value(wh[k]);
			sw += w[k];
		}
		out->dinv[I] = (float) (1.0 / sw);
	}
	free(rcnt);
	free(rkey);
	free(rval);
	if (rc)
		ggb_galop_free(out);
	return rc ? -1 : 0;
}
