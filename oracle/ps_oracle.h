/*
 * ps_oracle.h -- CPU restatement of MSBG's adaptive pressure relax
 * (block Gauss-Seidel with red-black block order, Laplacian multiply,
 * dense-level Jacobi) plus the multigrid pieces the reference names but does
 * not ship (set-up of flags/diagonal/face areas, constant restriction and
 * prolongation, the V-cycle).
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker, not the product: only the
 * test suite, the smoke check and the benchmark's CPU-baseline leg may load
 * it. The product (libmsbg_gpu.so, msbg_ps_*) never calls into it.
 *
 * Reference (paths relative to the reference's src/):
 *   relax()                      msbg3.cpp:2244-2369
 *   relaxBlockList()             msbg3.cpp:2170-2239 (red-black block order)
 *   IprocessBlockLaplacian()     msbg3.cpp:418-1600 (halo, sub-iterations)
 *   procNeighMixedRes2()         msbg3.cpp:59-160 (coarse-fine faces)
 *   multiplyLaplacianMatrixOpt() msbg3.cpp:1756-1912
 *   ImultiplyLaplacianMatrixDenseLevel / ...ProcessSliceDenseLevel
 *                                msbg3.cpp:1917-2165, 166-407
 *   configuration                msbgcell.h:97-101 (RELAX_BLOCK_GAUSS_SEIDEL,
 *                                RELAX_BLOCKS_RED_BLACK, MG_CONST_*),
 *                                msbg.cpp:819,837 (omega 6/7, 5 block iters)
 *
 * Channel layout per multigrid level m:
 *   block levels (m < nlev): the pages of all blocks in block-id order, block b
 *     at level L_b(m) = max(level_b, m) with bsx0 >> L_b(m) cells per edge,
 *     vid = (vz*bsx + vy)*bsx + vx.
 *   dense levels (m >= nlev): one array of (sx+2)*(sy+2)*(sz+2) floats with a
 *     one-cell zero ghost shell, sx = nbx*bsx0 >> m, x fastest.
 */
#ifndef PS_ORACLE_H
#define PS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* input cell types of the finest grid */
enum { PS_FLUID = 0, PS_SOLID = 1, PS_VOID = 2 };
/* cell flags, values of msbg.h:602-615 */
enum { PS_CELL_SOLID = 1 << 0, PS_CELL_PARTIAL_SOLID = 1 << 6, PS_CELL_VOID = 1 << 12 };
/* block flags, values of msbg.h:565,575 */
enum { PS_BLK_NO_FLUID = 1 << 5, PS_BLK_CUTS_SOLID = 1 << 15 };
/* domain boundary types, msbg.h:191-192 */
enum { PS_DBC_OPEN = 0, PS_DBC_SOLID = 1 };
/* options, values of msbg.h:115,121 */
enum { PS_OPT_CALC_RESIDUAL = 1 << 12, PS_OPT_REVERSE_ORDER = 1 << 18 };
/* channels */
enum { PS_CH_X = 0, PS_CH_B = 1, PS_CH_R = 2, PS_CH_T = 3, PS_CH_D = 4 };

typedef struct or_ps or_ps;

/* dbc[dir*2+side]; level[nb] block levels (face neighbours differ by <= 1);
 * cell_type[(nbx*bsx0)*(nby*bsx0)*(nbz*bsx0)], x fastest. Returns NULL and
 * *err != 0 on invalid input. */
or_ps *or_ps_create(int nbx, int nby, int nbz, int bsx0, double dx0, const int *dbc,
                    float omega, int block_iter, int ndense, const int8_t *level,
                    const uint8_t *cell_type, int *err);
void or_ps_destroy(or_ps *ps);
int or_ps_num_levels(const or_ps *ps, int *nmg, int *nlev);
int64_t or_ps_level_size(const or_ps *ps, int m);
int or_ps_get(const or_ps *ps, int chan, int m, float *out);
int or_ps_set(or_ps *ps, int chan, int m, const float *in);
int or_ps_get_flags(const or_ps *ps, int m, uint16_t *out);
int or_ps_get_block_flags(const or_ps *ps, int m, uint16_t *out);

int or_ps_relax(or_ps *ps, int reverse, int m, int chan_x, int chan_b, int chan_tmp, int n_iter);
int or_ps_multiply(or_ps *ps, unsigned options, int m, int chan_x, int chan_b, int chan_dst,
                   double *dot);
int or_ps_restrict(or_ps *ps, int m, int chan_src, int chan_dst);
int or_ps_prolongate(or_ps *ps, int m, int chan_src, int chan_dst);
int or_ps_fill(or_ps *ps, int chan, int m, float v);
int or_ps_vcycle(or_ps *ps, int nu1, int nu2, int coarse_iter);

#ifdef __cplusplus
}
#endif
#endif
