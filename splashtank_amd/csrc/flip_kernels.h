// flip_kernels.h - launch wrappers of the FLIP substep kernels (host side).
#pragma once
#include <hip/hip_runtime.h>

#include "flip_device.h"

// Everything one stage launch needs: the step's parameter block and collider
// records (device copies), the sim's device pointers, and launch sizes.
// entries of the per-step foam table (density sums 0 .. FOAM_TAB - 1)
#define FOAM_TAB 1024

// resident SOR: workgroup size, tiles resident per workgroup, and the largest
// grid its barrier serves (sorBar holds COOP_MAX_WG epochs, then the sequence
// number of the last solve it finished, a count of barrier time-outs, the
// barrier's release word and the path of the last solve: 0 one launch per
// iteration, 1 resident, 2 streaming)
#define COOP_WG 1024
#define COOP_TMAX 54
#define COOP_MAX_WG 1024

struct LaunchCtx {
  const DParams* P;
  const DCols* C;
  DGrid G;
  uint32_t n;         // particles (host copy of P->n)
  uint32_t ntiles;
  uint32_t scap;
  uint32_t cellGrid;  // workgroups of grid-stride loops over active cells
  uint32_t tileGrid;  // workgroups of loops with one workgroup per tile in use
  uint32_t pipeGrid;  // persistent SOR workgroups (about eight per CU)
  uint32_t coopGrid;  // resident-SOR workgroups (one per CU), 0 when cooperative launches are unavailable
  uint32_t sorSeq;    // sequence number of this pressure solve
  bool sorStream;     // the cooperative solve streams tile windows even when the tiles fit in LDS
  bool binned;        // blocks: need, binCount and binKey already hold A's marks and bins
  bool binNext;       // g2p: mark and bin the positions it writes, for the next substep
  bool nbOn;          // narrow band on this step
  uint32_t initStep;  // narrow-band start: 1 restart, 2 water added, 0 none
  uint32_t smax;      // whitewater slots (host copy of P->smax)
  uint32_t extrap;    // opt-in face velocity extrapolation: layers (host copy of P->extrap)
  uint8_t* exMask;    // its face masks, two per pool cell (ping-pong)
};

// scratch of the narrow-band reseed: dead-slot compaction over particles and the
// deficit scan over active cells (ncap = pool cells)
struct NbScratch {
  uint32_t* deadFlag;   // N
  uint32_t* deadStart;  // N + 1
  uint32_t* deadList;   // N
  uint32_t* need;       // ncap
  uint32_t* needStart;  // ncap + 1
  uint32_t* ent;        // ncap
  uint32_t* entIdx;     // ncap + 1
  uint32_t* sums;       // per-1024 block sums of the scans
  uint32_t ncap, nmax;
};

// Batched tanks.  Per tank: the device copies of the step's parameter block and
// collider records, and the tank's device pointers.
struct BTank {
  const DParams* P;
  const DCols* C;
  DGrid G;
};
#define ST_BATCH_MAX 256
// per tank, what one batched substep launches for it
struct BTankPlan {
  uint32_t n;       // particles (0: the tank sits this substep out)
  uint32_t ntiles;
  uint32_t scap;
  uint32_t iters;   // SOR iterations
  bool binned;      // need, binCount and binKey already hold A's marks and bins
  bool solid;       // rebuild the per-cell collider lookup
};
// Block tables of a batched substep: table k holds nt + 1 workgroup starts, one
// range per tank (empty where the tank has no work in that launch).  Table
// BT_PIPE0 + it belongs to SOR iteration it.
enum BatchTable {
  BT_FILL_MARK, BT_MARK, BT_NWG, BT_ONE, BT_ALLOC, BT_CELL, BT_NBR, BT_SCAN, BT_PART,
  BT_TILE, BT_ONE_N, BT_SOLID, BT_COPY, BT_FOAM, BT_G2P, BT_FILL_NEXT, BT_PIPE0
};
size_t batch_table_words(uint32_t nt, uint32_t itersMax);
// fills the tables (host copy); sorResident = SOR workgroups resident at once
void batch_tables(const BTankPlan* plan, uint32_t nt, uint32_t itersMax, uint32_t sorResident, uint32_t* tab);
// every stage of one substep for all tanks, one launch per kernel: dT and dtab
// are the device copies of the tank records and tables, htab the host copy
// coopGrid > 0 (and a plan buffer of 3 nt + 1 words): the SOR of every tank in
// one cooperative launch of that many resident workgroups, else (or when the
// launch is refused) one launch per iteration; returns whether it ran resident
bool launch_batch_step(hipStream_t st, const BTank* dT, const uint32_t* dtab, const uint32_t* htab, uint32_t nt, uint32_t itersMax,
                       uint32_t coopGrid, uint32_t* dPlan);

void launch_fill_u32(hipStream_t st, uint32_t* p, size_t n, uint32_t v);
void launch_blocks(hipStream_t st, const LaunchCtx& L);
void launch_clear(hipStream_t st, const LaunchCtx& L);
void launch_p2g(hipStream_t st, const LaunchCtx& L);
void launch_classify(hipStream_t st, const LaunchCtx& L, bool solid);
void launch_narrowband(hipStream_t st, const LaunchCtx& L);
void launch_reseed(hipStream_t st, const LaunchCtx& L, const NbScratch& S);
void launch_norm_faces(hipStream_t st, const LaunchCtx& L);
void launch_divergence(hipStream_t st, const LaunchCtx& L);
// returns the error of the resident-SOR cooperative launch (the per-iteration
// launches still solve when it failed)
hipError_t launch_pressure(hipStream_t st, const LaunchCtx& L, int iters);
// workgroups of the persistent SOR launch that fit on `cus` CUs at once, a
// multiple of 8 (the XCD count)
uint32_t sor_resident_grid(uint32_t cus);
// workgroups of the resident SOR (one per CU when it fits there), 0 if it does not
uint32_t sor_coop_grid(uint32_t cus);
// tile indices in Morton order of their (ti, tj, tk) coordinates
void morton_tiles(uint32_t tx, uint32_t ty, uint32_t tz, uint32_t* out);
void launch_project(hipStream_t st, const LaunchCtx& L);
void launch_g2p(hipStream_t st, const LaunchCtx& L);
void launch_unsort(hipStream_t st, const float4* src, const uint32_t* pid, float4* out, uint32_t n);
// the live view: what one packing needs (positions quantised to 16 bits across the box)
struct ViewParams {
  float x0, y0, z0, kx, ky, kz, vs, foamLife;  // kx = 65535 / (x1 - x0), ...
  uint32_t nW, nS, srcStride, fromPos;          // water points, whitewater capacity, water stride
  uint32_t oWp, oWs, oWf, oSp, oSk;             // word offsets of the parts
  uint32_t n;                                   // particles
};
// gathers the water (id order, every V.srcStride-th) and whitewater (sdMax sample slots,
// living ones compacted) samples and packs them into `out` (V's word offsets); flag and
// off hold sdMax (+1) words, sums sdMax / 1024 + 2, sdisp sdMax points
// deep water points of the narrow band look (at most DEEP_MAX, the reference's
// DEEPMAX); flags, scan and emit over the
// ncell = slots * 512 cells of the pool; all[0] counts every deep cell, off[ncell]
// the kept ones (at most ST_DEEP_MAX written to out)
#define DEEP_MAX 200000u
void launch_deep(hipStream_t st, const LaunchCtx& L, uint32_t ncell, uint32_t prev, uint32_t* flag, uint32_t* off, uint32_t* sums,
                 uint32_t* all, float4* out);
void launch_pack_view(hipStream_t st, const LaunchCtx& L, const ViewParams& V, uint32_t sdMax, uint32_t* flag, uint32_t* off,
                      uint32_t* sums, float4* sdisp, uint32_t* out);
// the spray stage (wgsl-whitewater.js): the air field when `air` (airN cells),
// then every whitewater point, then new points when `emit`, then the drops of
// ndrip drip samples (0: none)
void launch_spray(hipStream_t st, const LaunchCtx& L, uint32_t airN, bool air, bool emit, uint32_t ndrip);
// FP64 norms of the last pressure solve and projection (flip_norms.hip): part
// holds L.cellGrid x ST_NORM_WORDS doubles, out ST_NORM_WORDS (the st_norms fields)
#define ST_NORM_WORDS 8
void launch_norms(hipStream_t st, const LaunchCtx& L, double* part, double* out);
// the fixed scratch of the allocation-free norms on device `dev` (which must be
// current): room for the partials of NORMS_STATIC_WG workgroups and the results
#define NORMS_STATIC_WG 4096u
#define NORMS_STATIC_DEVICES 64
hipError_t norms_static_scratch(int dev, double** ptr, size_t* words);
