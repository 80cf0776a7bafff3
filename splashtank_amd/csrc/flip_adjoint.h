// flip_adjoint.h - adjoint (vector-Jacobian product) of the FLIP substep: the
// recorded state of one substep and the launch of its backward pass.
#pragma once
#include <hip/hip_runtime.h>

#include "flip_device.h"

// What the backward pass of one substep reads, recorded while the substep ran
// forward.  Per-slot arrays cover the whole pool (scap * 512 cells).
struct AdjTape {
  float4* pos;          // particle state at the start of the substep (its A order)
  float4* vel;
  uint32_t* pid;
  uint32_t* binStart;   // the substep's particle bins by tile (ntiles + 2 starts into binIdx)
  uint32_t* binIdx;
  uint32_t* tmap0;      // tile map at the start of the substep
  uint32_t* tmap;       // tile map, slot tiles and slot list after blocks
  uint32_t* slotTile;
  uint32_t* slotList;
  uint32_t* ctl;        // control words at the end of the substep (ctl[4] = tiles in use)
  uint32_t* ctype;      // cell types after classify
  uint8_t* sten;        // SOR stencil byte (bit 7 fluid, bits 0-2 non-solid neighbours)
  int32_t* accWraw;     // fixed-point weight sums after p2g (3 per cell)
  int32_t* accVfin;     // new face velocity after project (f32 bits)
  int32_t* accWfin;     // old face velocity (f32 bits)
  DParams* P;           // the substep's parameter block (device copy)
  DCols* C;             // the substep's collider records (device copy)
  const float* sdf;     // collider distance grids (the sim's, unchanged since the record)
  uint32_t n;           // particles
  uint32_t iters;       // SOR iterations
};

// Gradient buffers of one adjoint context (device).
struct AdjGrad {
  float4* gPos;   // dL/d(particle position), id order (N)
  float4* gVel;   // dL/d(particle velocity), id order (N)
  float4* gPos2;  // the same for the state before the substep being reversed
  float4* gVel2;
  float* gFN;     // per face: dL/d(new face velocity), then dL/d(velocity sum)
  float* gFO;     // per face: dL/d(old face velocity), then dL/d(weight sum)
  float* gPres;   // per cell: dL/d(pressure), SOR ping-pong with gPres2
  float* gPres2;
  float* gRhs;    // per cell: dL/d(rhs)
  float* carry;   // per tile of the box (ntiles * 512): dL/d(pressure at the end of the
                  // previous substep), through the next substep's warm start
  uint32_t* sorNb;  // per cell, 6 per cell: the fluid neighbour a transposed sweep gathers from
                    // (x-, x+, y-, y+, z-, z+), ST_NONE where there is none or it has no non-solid side
  float* sorCoef;   // per cell, 6 per cell: omega / (non-solid neighbours of that neighbour)
  float* gCol;      // per collider of the substep reversed last, 9 words: dL/d(centre),
                    // dL/d(linear velocity), dL/d(angular velocity)
  uint32_t* bar;    // ADJ_BAR_WORDS: the one-launch solve's grid barrier (flags, time-outs, release)
  uint32_t ncell; // scap * 512
  uint32_t ncarry;
};

// colliders the adjoint follows (pushes of one particle are kept in registers)
#define ADJ_MAX_COLS 4

// workgroups of the one-launch transposed solve: at most ADJ_SOR_MAX_WG, so
// that workgroup 0 waits for each barrier flag with one thread; the barrier
// words are the flags, then the time-out count (ADJ_BAR_FAILS) and the release
// word (ADJ_BAR_RELEASE)
#define ADJ_SOR_MAX_WG 256u
#define ADJ_BAR_FAILS (ADJ_SOR_MAX_WG + 1u)
#define ADJ_BAR_RELEASE (ADJ_SOR_MAX_WG + 2u)
#define ADJ_BAR_WORDS (ADJ_SOR_MAX_WG + 4u)

// resident workgroups of the one-launch transposed solve on the current device
// (0 when cooperative launches are not available)
uint32_t adj_sor_coop_grid(uint32_t cus);

// the backward pass of one recorded substep: reads G.gPos/gVel (gradient of the
// state after it), writes G.gPos2/gVel2 (gradient of the state before it) and the
// pressure carry for the substep before.  With coopGrid > 0 the transposed
// solve runs as one cooperative launch of that many workgroups (the same
// arithmetic as one launch per sweep, which runs when coopGrid is 0 or the
// cooperative launch is refused); returns whether the one launch ran.
bool launch_adjoint_step(hipStream_t st, const AdjTape& T, const AdjGrad& G, uint32_t scap, uint32_t ntiles, uint32_t coopGrid);
