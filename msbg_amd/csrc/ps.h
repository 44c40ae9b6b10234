// Device-side descriptors and launchers of the adaptive pressure relax
// (msbg3.cpp): block relax / Laplacian multiply on block levels, Jacobi and
// multiply on dense levels, constant restriction and prolongation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace msbg {

constexpr int kPsMaxBlockLevels = 4;  // block edge 16 -> 16, 8, 4, 2

// cell and block flags with the reference's values (msbg.h:565-615)
constexpr uint16_t kPsCellSolid = 1u << 0;
constexpr uint16_t kPsCellPartialSolid = 1u << 6;
constexpr uint16_t kPsCellVoid = 1u << 12;
constexpr uint16_t kPsBlkNoFluid = 1u << 5;
constexpr uint16_t kPsBlkCutsSolid = 1u << 15;

enum { PS_MODE_RELAX = 0, PS_MODE_MULTIPLY = 1, PS_MODE_RESIDUAL = 2 };

// One block multigrid level as the block kernel reads it.
struct PsBlockArgs {
  const float *X;           // values the halo is built from
  const float *B;           // right-hand side
  float *Y;                 // destination (the relax writes X in place)
  const float *D;           // diagonal coefficient
  const uint16_t *flags;    // cell flags
  const uint16_t *bflags;   // block flags
  const float *fa[kPsMaxBlockLevels][3];  // page-form face areas of every block, per level
  double dx0;
  float omega;
  int openRight[3];         // domain's upper border open (1) or solid (0), per axis
  int rev;                  // OPT_REVERSE_ORDER given
  int biter;                // sub-iterations per block (_mgSmBlockIter)
  double *dot;              // multiply: one partial dot product per workgroup
};

// One listed block with everything its halo needs, so that the kernel reads
// one record instead of chasing block tables.
struct alignas(16) PsItem {
  int64_t off;      // page offset of the block (floats)
  int64_t noff[6];  // page offsets of the face neighbours (0 outside the domain)
  int32_t b;        // block id
  int32_t nbr[6];   // face neighbours (-x,+x,-y,+y,-z,+z), -1 outside the domain
  int8_t L;         // the block's level at this multigrid level
  int8_t nL[6];     // the neighbours' levels
  int8_t pad;
};

// One dense multigrid level (padded by a one-cell zero shell).
struct PsDenseArgs {
  int sx, sy, sz;
  const float *X, *B;
  float *Y;
  const float *D;
  const uint16_t *flags;
  const float *fa[3];
  float h, hromega, om1;
};

// Transfer between mg level m (fine) and m+1 (coarse).
struct PsXferArgs {
  int kind;               // 0 block->block, 1 block->dense, 2 dense->dense
  int nbx, nby, nbz, bsx0;
  const int8_t *Lf, *Lc;  // block levels at m and m+1 (kinds 0, 1)
  const int64_t *offF, *offC;
  int fsx, fsy, fsz;      // dense fine dims (kind 2)
  int csx, csy, csz;      // dense coarse dims (kinds 1, 2)
  const uint16_t *flagsF, *flagsC;
};

// cuts: every listed block has BLK_CUTS_SOLID (face weights), or none has
int launch_ps_block(hipStream_t s, int bsx, int mode, bool cuts, const PsBlockArgs &a,
                    const PsItem *items, int n, int *nWorkgroups);
int ps_block_workgroups(int bsx, int n);
int launch_ps_dense(hipStream_t s, int mode, const PsDenseArgs &a);
int launch_ps_restrict(hipStream_t s, const PsXferArgs &a, const float *src, float *dst);
int launch_ps_prolongate(hipStream_t s, const PsXferArgs &a, const float *src, float *dst);
int launch_ps_fill(hipStream_t s, float *p, int64_t n, float v);
int launch_ps_copy(hipStream_t s, const float *src, float *dst, int64_t n);

}  // namespace msbg
