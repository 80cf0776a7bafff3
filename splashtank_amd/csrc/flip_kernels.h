// flip_kernels.h - launch wrappers of the FLIP substep kernels (host side).
#pragma once
#include <hip/hip_runtime.h>

#include "flip_device.h"

// Everything one stage launch needs: the step's parameter block and collider
// records (device copies), the sim's device pointers, and launch sizes.
struct LaunchCtx {
  const DParams* P;
  const DCols* C;
  DGrid G;
  uint32_t n;         // particles (host copy of P->n)
  uint32_t ntiles;
  uint32_t scap;
  uint32_t cellGrid;  // workgroups of grid-stride loops over active cells
  uint32_t tileGrid;  // workgroups of loops with one workgroup per tile in use
  uint32_t sorGrid;   // workgroups of the SOR launches (one tile each when it fits)
  uint32_t pipeGrid;  // persistent SOR workgroups (about eight per CU)
};

void launch_fill_u32(hipStream_t st, uint32_t* p, size_t n, uint32_t v);
void launch_blocks(hipStream_t st, const LaunchCtx& L);
void launch_clear(hipStream_t st, const LaunchCtx& L);
void launch_p2g(hipStream_t st, const LaunchCtx& L);
void launch_classify(hipStream_t st, const LaunchCtx& L, bool solid);
void launch_norm_faces(hipStream_t st, const LaunchCtx& L);
void launch_divergence(hipStream_t st, const LaunchCtx& L);
void launch_pressure(hipStream_t st, const LaunchCtx& L, int iters);
void launch_project(hipStream_t st, const LaunchCtx& L);
void launch_g2p(hipStream_t st, const LaunchCtx& L);
void launch_unsort(hipStream_t st, const float4* src, const uint32_t* pid, float4* out, uint32_t n);
