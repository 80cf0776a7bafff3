/*
 * splashtank.h - C ABI of the MI355X FLIP substep on the sparse 8^3-block MAC grid.
 *
 * One `st_sim` owns all device memory of one tank (particles, tile map, block
 * pool) and one HIP stream.  Every call returns ST_OK (0) or a negative ST_ERR_*
 * code; st_last_error() gives the message of the calling thread's last failure.
 * Nothing here takes torch or JS types: plain pointers and sizes only.
 *
 * The functions replace the WebGPU-side members of the reference's JS class
 * `FlipSim` (src/flip/sim.js:46-547); the citation on each line names the member
 * it stands in for.  The N-API addon (splashtank_amd/csrc/napi_addon.cpp) and the
 * Python ctypes loader (splashtank_amd/capi.py) bind exactly these symbols.
 */
#ifndef SPLASHTANK_H
#define SPLASHTANK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ST_OK 0
#define ST_ERR_ARG (-1)   /* bad argument (null, size, range) */
#define ST_ERR_HIP (-2)   /* a HIP runtime call failed */
#define ST_ERR_STATE (-3) /* call not valid in this state */
#define ST_ERR_NOMEM (-4) /* device allocation failed */

/* open faces of the box (sim.js:55; ctypeOf at wgsl-core.js:104-112) */
#define ST_OPEN_X0 1u
#define ST_OPEN_X1 2u
#define ST_OPEN_Y0 4u
#define ST_OPEN_Y1 8u
#define ST_OPEN_Z0 16u
#define ST_OPEN_Z1 32u

/* Creation options: `new FlipSim(device, opt)` (sim.js:47-65). */
typedef struct st_opts {
  double d;        /* particle spacing (opt.d) */
  double ratio;    /* cell size / spacing, clamped to [1.4, 4]; 0 -> 2 (opt.ratio) */
  double size[3];  /* box extent W, H, D; 0 -> (3, 2.6, 3) (opt.size) */
  uint32_t open;   /* ST_OPEN_* bitmask (opt.open) */
  uint32_t max_n;  /* particle capacity (opt.maxN) */
  uint32_t slots;  /* block pool size incl. spare slot 0; 0 -> every tile + 1 (opt.slots) */
  int32_t device;  /* HIP device ordinal */
} st_opts;

/* Derived sizes, the public fields of FlipSim (sim.js:48-65). */
typedef struct st_dims {
  float d, h, ratio, rest;
  float x0, y0, z0, x1, y1, z1; /* box: centred in x and z, y from 0 (sim.js:53) */
  int32_t nx, ny, nz;           /* cells per axis, round(size/h) >= 4 (sim.js:51) */
  int32_t tx, ty, tz;           /* tiles per axis, ceil((n+1)/8) (sim.js:60) */
  uint32_t ntiles, scap, scap_max, max_n, ncell;
} st_dims;

/* Per-step settings: the `o` argument of FlipSim.step(dt, o) (sim.js:179-241)
 * and the FlipSim.opt flags (sim.js:49).  st_params_default() fills the
 * reference's defaults; zero in `omega` or `vmax` means the default as in
 * `o.omega || this.omega()` and `o.vmax || 14` (sim.js:197-198). */
typedef struct st_params {
  float dt;
  float g;          /* gravity magnitude, subtracted from v.y (o.g) */
  float flip;       /* FLIP ratio (o.flip, 0.95) */
  float comp;       /* density correction gain (o.comp, 6) */
  float omega;      /* SOR over-relaxation (o.omega, default min(1.9, 2/(1+sin(pi/max n)))) */
  float vmax;       /* particle speed cap (o.vmax, 14) */
  float tear;       /* sheet tearing threshold in particles per cell (o.tear, 0) */
  int32_t iters;    /* red-black SOR iterations (opt.iters || o.iters, at least 1) */
  float wind[3];    /* o.wind */
  float drag;       /* o.drag (0) */
  float foam_tau;   /* o.foamTau (1.2) */
  float foam_sens;  /* o.foamSens (1) */
  float gust;       /* o.gust (0) */
  uint32_t no_foam; /* opt.noFoam */
  uint32_t no_cull; /* opt.noCull */
  uint32_t fast_air;/* opt.fastAir */
  uint32_t no_solid_cache; /* opt.noSolidCache */
  uint32_t nb_on;   /* narrow band on (o.nb.on) */
  float nb_band;    /* narrow band width in cells, at least 2 (o.nb.band, 3) */
  uint32_t only_tiles; /* run the block stage only (FlipSim.only = 'tiles', sim.js:173) */
  /* narrow-band start (sim.js:188-220, nbSeed at shared.js:805-853): nb_init 1 restart,
   * 2 water added, 0 none; up to 8 fill shapes of 8 floats each, box
   * (1, x0, y0, z0, x1, y1, z1, 0) or sphere (2, x, y, z, r, 0, 0, 0); nfv trailing
   * entries of the collider set are water volumes, not colliders. */
  uint32_t nb_init;
  uint32_t nfill;
  uint32_t nfv;
  float fills[64];
  /* pressure solve: 0 (default) runs every iteration of a solve in one
   * cooperative launch, with the pressure resident in LDS when the tiles in use
   * fit and streaming tile windows otherwise; 1 one launch per red-black
   * iteration; 2 the cooperative launch, always streaming; 3 (ST_SOR_TUNED) the
   * path and launch width st_tune_sor chose for this sim (0 until it ran).  All
   * give the same bits.  No counterpart in the reference (an implementation
   * choice here). */
  uint32_t sor_mode;
  /* whitewater (o.spr, o.spr.ww, o.air; sim.js:203-206, 221-230).  The spray
   * stage runs while the sim has room for whitewater points
   * (st_set_spray_cap): the points there move every substep, and new ones are
   * thrown off when spr_on. */
  uint32_t spr_on;      /* o.spr given */
  float spr_rate;       /* o.spr.rate: points per second per unit of emission */
  float spr_min;        /* o.spr.min: crest speed where spray starts (at least 0.05) */
  float spr_life;       /* o.spr.life: seconds a spray point lives */
  float spr_mist;       /* o.spr.mist: share of fast spray that is mist */
  float spr_drag;       /* o.spr.drag: spray drag toward the air, 1/s */
  float spr_mdrag;      /* o.spr.mdrag: mist drag, 1/s */
  float vs;             /* o.vs: metres per scene unit (0 -> 1) */
  float ww_foam_life;   /* o.spr.ww.foamLife, s (0 -> 3) */
  float ww_rise;        /* o.spr.ww.rise: bubble rise speed (0 -> 0.3) */
  float ww_crest;       /* o.spr.ww.crest: weight of crest emission (1) */
  float ww_impact;      /* o.spr.ww.impact: weight of churn emission (0.5) */
  uint32_t air_on;      /* o.air.on: a coarse air field the spray drags toward (with spr_on) */
  float air_couple;     /* o.air.couple, 1/s */
  float air_decay;      /* o.air.decay, 1/s */
  float air_blur;       /* o.air.blur (0.35) */
  uint32_t drip_on;     /* o.drip.on: wet colliders drip (with spr_on and drip samples set) */
  float drip_rate;      /* o.drip.rate: drops per second per sample when fully wet */
  float drip_dry;       /* o.drip.dry: seconds a sample takes to dry */
  /* ocean side boundary (o.ocean and o.time; sim.js:184-218, words 57-66 and
   * 148 of the parameter block).  ocean_on turns the narrow band on too, with
   * nb_band as its width (sim.js:185-186).  The wave list is the one
   * st_ocean_waves gives for the ocean's settings, gravity and depth
   * ocean_level - y0. */
  uint32_t ocean_on;
  uint32_t nwave;       /* waves in `waves` (at most ST_OCEAN_MAX_WAVES) */
  float ocean_time;     /* o.time: seconds since the start */
  float ocean_level;    /* o.ocean.level: height of the still water */
  float ocean_layer;    /* o.ocean.layer: width of the boundary band at the side walls */
  float waves[8 * 32];  /* per wave: kx, kz, amplitude, angular frequency, phase, wavenumber, 0, 0 */
  /* Opt-in modes the reference does not have.  0 (the default) is the
   * reference's path, bit for bit. */
  uint32_t rk2;           /* 1: a particle of a tile with storage moves by the midpoint rule
                           * through the grid's new velocity, p + dt u(p + dt/2 u(p)), instead
                           * of p + dt v (its velocity stays the FLIP/PIC blend) */
  uint32_t extrap_layers; /* face velocities (new and old) carried from the faces next to the
                           * fluid into the air around it after the projection, this many
                           * layers (0 off, at most ST_EXTRAP_MAX) */
} st_params;

#define ST_EXTRAP_MAX 4u

#define ST_OCEAN_MAX_WAVES 32u

/* The ocean's settings (o.ocean as the reference's page sets it: height, length, dir, spread, seed). */
typedef struct st_ocean {
  double height;  /* significant wave height */
  double length;  /* peak wavelength */
  double dir;     /* main direction, degrees */
  double spread;  /* directional spread, degrees */
  int32_t seed;
} st_ocean;
/* oceanWaves(oc, g, depth) (src/shared.js:769-791): 24 linear waves around the
 * peak wavelength with a Pierson-Moskowitz spectrum scaled to the significant
 * height, spread around the direction, each with the finite-depth dispersion
 * w^2 = g k tanh(k depth).  Writes 8 floats per wave (the layout of
 * st_params.waves) for at most `cap` waves and sets *n to 24.  Host code. */
int st_ocean_waves(const st_ocean* oc, double g, double depth, float* w, uint32_t cap, uint32_t* n);

#define ST_SOR_TUNED 3u

/* Stages of one substep, named as in FlipSim.stages (sim.js:245-310). */
enum {
  ST_STAGE_SORT = 0,
  ST_STAGE_BLOCKS = 1,
  ST_STAGE_CLEAR = 2,
  ST_STAGE_P2G = 3,
  ST_STAGE_NARROWBAND = 4,
  ST_STAGE_CLASSIFY = 5,
  ST_STAGE_NORMFACES = 6,
  ST_STAGE_DIVERGENCE = 7,
  ST_STAGE_PRESSURE = 8,
  ST_STAGE_PROJECT = 9,
  ST_STAGE_G2P = 10,
  ST_STAGE_SPRAY = 11,
  ST_STAGE_RESEED = 12,
  ST_STAGE_COUNT = 13
};

/* Canonical dumps: per-cell fields are given for every tile that holds storage,
 * tiles in increasing tile number t = ti + tx*(tj + ty*tk), 512 cells each in
 * local order i&7 + 8*((j&7) + 8*(k&7)), faces as 3 values per cell (the cell's
 * lower x, y, z faces).  Particles come in id order (sim.js:439-455). */
enum {
  ST_DUMP_TILES = 0,  /* u32 tile numbers with storage, ascending */
  ST_DUMP_CNT = 1,    /* i32 particles per cell */
  ST_DUMP_CTYPE = 2,  /* u32 type | (collider+1)<<8 | phi<<16 */
  ST_DUMP_PRES = 3,   /* f32 pressure */
  ST_DUMP_RHS = 4,    /* f32 right-hand side (narrow band: depth scratch) */
  ST_DUMP_ACCV = 5,   /* i32 x3 fixed-point velocity sums (valid after p2g, before normFaces) */
  ST_DUMP_ACCW = 6,   /* i32 x3 fixed-point weight sums (same window) */
  ST_DUMP_FACEV = 7,  /* f32 x3 new face velocity (after normFaces; projected after project) */
  ST_DUMP_FACEO = 8,  /* f32 x3 old face velocity (after normFaces) */
  ST_DUMP_VELG = 9,   /* f32 x3 kept grid velocity (narrow band) */
  ST_DUMP_POS = 10,   /* f32 x4 particle positions, id order */
  ST_DUMP_VEL = 11,   /* f32 x4 particle velocities, id order */
  ST_DUMP_TMAP = 12,  /* u32 tile -> slot or 0xffffffff, ntiles entries */
  ST_DUMP_COUNT = 13
};

/* thread-local message of the last failed call on this thread */
const char* st_last_error(void);
/* library version string */
const char* st_version(void);

/* reference defaults of FlipSim.step's `o` plus FlipSim.opt (sim.js:49, 197-199) */
void st_params_default(st_params* p);

/* new FlipSim(device, opt): buffers, empty tile map, free stack scap-1..1 (sim.js:47-105) */
int st_create(const st_opts* opts, void** sim_out);
/* FlipSim.destroy (sim.js:496-501) */
int st_destroy(void* sim);
/* derived sizes (sim.js:48-65) */
int st_get_dims(void* sim, st_dims* out);

/* GpuSim.flushRun: write particles [first, first+count) from host float4
 * arrays; N grows to first+count (gpu.js:390-408).  pos.w and vel.w are kept. */
int st_upload_particles(void* sim, uint32_t first, uint32_t count, const float* pos4, const float* vel4);
/* GpuSim.setCap / clear: set the live particle count (gpu.js:388-389, sim.js:539-547) */
int st_set_count(void* sim, uint32_t n);
/* FlipSim.clear: N = 0, ids back to slot numbers, time and seed restart (sim.js:539-547) */
int st_clear(void* sim);
/* live particle count N */
int st_get_count(void* sim, uint32_t* n);

/* packColliders (gpu.js:42-70): nc colliders of 32 floats each, then all distance
 * grids back to back (nsdf floats).  `sig` identifies the grid set; the grids are
 * uploaded only when it changes.  A change of any collider float marks the
 * per-cell collider lookup dirty, as in sim.js:191-195. */
int st_set_colliders(void* sim, const float* cf32, uint32_t nc, const float* sdf, size_t nsdf, uint64_t sig);

/* FlipSim.step(dt, o): one full substep enqueued on the sim's stream (sim.js:179-241). */
int st_step(void* sim, const st_params* p);
/* Run several substeps with the same settings, as many st_step calls; used by
 * benchmarks and bakes.  With use_graph != 0 each substep is a HIP graph
 * replay: the first substep of a kind (particle buffer parity, bins made or
 * not, collider lookup rebuilt or not, sizes) is stream-captured and later
 * ones replay it, with the step's parameters copied into the one device slot
 * the graph reads.  Graphed substeps solve with one launch per SOR iteration
 * (the same bits; a captured cooperative launch is not scheduled as one when
 * other streams run at the same time).  Graphs cover plain substeps only
 * (ST_ERR_ARG with the narrow band or a tile-only step).
 * No counterpart in the reference, which records its dispatches into one
 * compute pass per step (src/flip/sim.js:245-310). */
int st_step_n(void* sim, const st_params* p, int32_t nsteps, int32_t use_graph);
/* The parameter upload of step() without running a stage (stepTimed, sim.js:312-319). */
int st_begin_step(void* sim, const st_params* p);
/* Enqueue one stage of the substep begun by st_begin_step (stepTimed, sim.js:320-328). */
int st_stage(void* sim, int32_t stage);

/* stepTimed (sim.js:312-331): one full substep as st_step, with a HIP event pair
 * around every stage on the sim's stream; ms[k] = elapsed milliseconds of stage k
 * (0 for stages that launch nothing).  Blocks until the substep is done. */
int st_time_stages(void* sim, const st_params* p, float ms[ST_STAGE_COUNT]);

/* FlipSim.readbackInto: N float4 positions in id order (sim.js:439-455); blocks. */
int st_read_particles(void* sim, float* dst4, uint32_t n);
/* readStats: copy of the 16 control words (sim.js:160-168): 0,1 red/black fluid
 * cells, 2 free stack top, 3 new tiles, 4 tiles in use, 5 tiles with no room,
 * 6 fastest particle speed (f32 bits). */
int st_read_ctl(void* sim, uint32_t ctl[16]);
/* FlipSim.grow(n): enlarge the block pool to n slots keeping contents (sim.js:137-152). */
int st_grow(void* sim, uint32_t slots);
/* Bytes needed by st_dump for `what` in the current state. */
int st_dump_size(void* sim, int32_t what, size_t* bytes);
/* Canonical dump of a field (see ST_DUMP_*); blocks until the stream is idle. */
int st_dump(void* sim, int32_t what, void* dst, size_t bytes);
/* Wait for all work of the sim (queue.onSubmittedWorkDone). */
int st_sync(void* sim);
/* Device memory held by the sim in bytes (memBytes, sim.js:492-495). */
int st_mem_bytes(void* sim, uint64_t* bytes);
/* Which path the last pressure solve took (waits for the stream): 0 one launch
 * per red-black iteration, 1 one cooperative launch with the pressure resident
 * in LDS, 2 one cooperative launch streaming tile windows (see
 * st_params.sor_mode).  No counterpart in the reference; for benchmarks and
 * tests. */
int st_sor_path(void* sim, uint32_t* path);

/* setSprayCap(n) (sim.js:333-340): room for n whitewater points (0 turns the
 * spray stage off).  The new pool holds no points; the ring counter of handed
 * out slots carries on. */
int st_set_spray_cap(void* sim, uint32_t n);
/* dripUpdate's upload (sim.js:353-363): n drip samples of 8 floats each -- the
 * local position on a collider, the collider's number, the local normal, the
 * wetness (the reference's dripSamples, src/flip/bake.js:19-40, gives them dry)
 * -- replacing the ones before (n = 0 removes them).  The spray stage lets the
 * samples drip when st_params.drip_on. */
int st_set_drips(void* sim, const float* samples, uint32_t n);
/* readSpray (sim.js:423-437): the first n = max(1, min(cap, slots handed out))
 * whitewater slots, positions with the seconds of life left in w and
 * velocities with the kind in w (0 spray, 1 mist, 2 foam, 3 bubble; the
 * fraction is how long the point has had that kind); *n = 0 without a pool.
 * Writes at most `room` points; blocks. */
int st_read_spray(void* sim, float* pos4, float* vel4, uint32_t room, uint32_t* n);

/* The live view (dispPack + pump, sim.js:364-404, 504-537; packDisp and
 * sprayDisp, wgsl-whitewater.js:183-253): a sample of the water (every stride-th
 * particle in id order, stride = max(1, ceil(N / disp_max)), disp_max 0 ->
 * 300000) and of the whitewater (up to min(pool, 200000) sample slots, the
 * living ones packed together in slot order) packed on the GPU into 32-bit
 * words in the cache frame format: a 4-word header (whitewater points, deep
 * points = 0, 0, 0), then the parts at the layout's word offsets.  With out ==
 * NULL only the layout is filled; else out must hold layout->words words.
 * Deep-water look points (setDeepLook) are not built (nD = 0).  Blocks. */
typedef struct st_view_layout {
  uint32_t nW, nS, nD;          /* water points, whitewater capacity, deep points */
  uint32_t words;               /* words of the packed view */
  uint32_t oWp, oWs, oWf;       /* water: positions (6 words per 4 points), speed bytes, white-water bytes */
  uint32_t oSp, oSk;            /* whitewater: positions, 32-bit codes */
  uint32_t oDp, oDs;            /* deep water: positions, speed bytes */
  uint32_t stride, from_pos;    /* water sample stride; 1 when the water is read straight from the particles */
} st_view_layout;
int st_pack_view(void* sim, uint32_t disp_max, uint32_t* out, size_t cap_words, st_view_layout* layout);

/* Deep water points of the narrow band look (setDeepLook/readDeep, sim.js:405-419;
 * deepClear/deepPick/deepFin, wgsl-whitewater.js:279-297): the centre of every fluid
 * cell in the domain that holds no particle after the last substep, as (x, y, z,
 * 4096 + min(|kept grid velocity|, 63)).  With more than ST_DEEP_MAX such cells at
 * the previous call, an even hash-chosen sample of them (the reference's rule).
 * Points come in active-cell order (the reference's come in arrival order; compare
 * as sets).  *n = 0 when the narrow band is off or the sim is empty.  At most
 * min(room, ST_DEEP_MAX) points are written; *n is how many were. */
#define ST_DEEP_MAX 200000u
int st_read_deep(void* sim, float* out4, uint32_t room, uint32_t* n);

/* Asynchronous readback, for a caller that must not block: replaces the
 * reference's mapAsync readbacks -- readbackInto (src/flip/sim.js:439-455),
 * readStats (sim.js:160-168), readSpray (sim.js:423-437) and readDeep
 * (sim.js:405-419) -- behind the N-API addon's Promises.
 * st_read_begin enqueues on the sim's stream, behind all work enqueued before
 * it, the copy of one result into pinned host memory that the ticket owns, and
 * returns without waiting.  st_read_wait may be called from any thread: it
 * waits for that copy alone (an event) and copies the result out, touching
 * nothing of the sim, so the sim may step on meanwhile (the result is the state
 * at st_read_begin).  st_read_free releases the ticket (any thread; it waits for
 * the copy first).  A ticket outlives its sim.
 *   ST_READ_PARTICLES: min(n, count) particles in id order (x, y, z, w) into
 *     out0 (room >= that many); *n = how many.
 *   ST_READ_CTL: the 16 control words (st_read_ctl) into out0 (room >= 16).
 *   ST_READ_SPRAY: as st_read_spray, positions into out0 and velocities into
 *     out1 (room points each); *n = 0 without a pool.
 *   ST_READ_DEEP: as st_read_deep into out0 (room points); *n = 0 when it
 *     gives none.  The count that picks the next read's sample is taken when
 *     st_read_wait runs. */
#define ST_READ_PARTICLES 0u
#define ST_READ_CTL 1u
#define ST_READ_SPRAY 2u
#define ST_READ_DEEP 3u
int st_read_begin(void* sim, uint32_t what, uint32_t n, void** ticket);
int st_read_wait(void* ticket, void* out0, void* out1, uint32_t room, uint32_t* n);
int st_read_free(void* ticket);

/* buildSDF(mesh, cell, margin) (src/shared.js:98-181): the signed distance grid
 * of a closed triangle mesh, the collider grids st_set_colliders takes.  v holds
 * nv vertices (x, y, z doubles), tri nt triangles (vertex indices; the
 * reference's meshTris fan triangulation of its faces).  cell <= 0 -> 0.025,
 * margin < 0 -> 0.16; the cell grows by 1.25 until the grid has at most 2.6e6
 * points.  Fills *grid (dims, origin = mesh bounds minus the margin, cell); with
 * data != NULL also the grid (x fastest, negative inside), which needs cap >=
 * dims[0]*dims[1]*dims[2] floats.  Host code: touches no GPU. */
typedef struct st_sdf_grid {
  int32_t dims[3];
  double origin[3];
  double cell;
} st_sdf_grid;
int st_build_sdf(const double* v, uint32_t nv, const uint32_t* tri, uint32_t nt, double cell, double margin, st_sdf_grid* grid,
                 float* data, size_t cap);

/* FP64 norms of the last substep's pressure solve and projection, over its fluid
 * cells, computed on the GPU in f64 from the f32 state the substep left (the
 * only f64 arithmetic on the GPU).  No counterpart in the reference, which
 * computes no residual (its solve runs a fixed count of red-black iterations,
 * src/flip/wgsl-solver.js:124-164).  With r = rhs + sum of fluid-neighbour
 * pressures - nns * p (the SOR system of relax), div = sum of (v_hi - v_lo) / h
 * over the projected faces, and tgt the density target of the divergence stage
 * (wgsl-solver.js:114-121), the projection makes div - tgt = -(dt/h^2) r up to
 * f32 rounding. */
typedef struct st_norms {
  double res_l2, res_max;  /* ||r||_2, max |r| */
  double rhs_l2;           /* ||rhs||_2 (res_l2 / rhs_l2 = relative residual) */
  double div_l2, div_max;  /* projected divergence */
  double err_l2, err_max;  /* div - tgt */
  double fluid;            /* fluid cells counted */
} st_norms;
/* Enqueues the two reduction kernels on the sim's stream after its last
 * substep and waits for them.  ST_ERR_STATE before the first substep. */
int st_solve_norms(void* sim, st_norms* out);
/* st_solve_norms without any device allocation.  No counterpart in the
 * reference; for callers that cannot allocate device memory (serving loops and
 * harnesses that forbid hipMalloc).  Its per-workgroup partial sums and results
 * live in one fixed array of 4097 x 8 doubles (256 KiB) per GPU, looked up on
 * the first call on that GPU and held until the process ends; every call on a
 * GPU shares it, so calls that overlap on one GPU race.  ST_ERR_ARG when the
 * sim's launch needs more than the array holds (its reductions use at most
 * 4096 workgroups).  The same results as st_solve_norms. */
int st_solve_norms_static(void* sim, st_norms* out);

/* Autotuner of the pressure-solve schedule.  No counterpart in the reference
 * (its solve has one fixed dispatch pattern, src/flip/sim.js:291-293).  Every
 * candidate gives the same bits, so tuning changes only speed.
 * st_tune_sor: from the sim's current state, runs one substep's stages up to
 * the divergence with `p` (narrow band off), then times the pressure stage for
 * each candidate -- the resident cooperative solve, the streaming cooperative
 * solve, and one launch per iteration at 1, 1/2 and 1/4 of the resident
 * workgroup count -- `reps` times each after one warm-up, keeps the fastest for
 * steps with sor_mode = ST_SOR_TUNED, and restores the state it started from
 * (the next substep gives the bits it would have given without the tuning).
 * st_batch_tune_sor: the same for a batch: it times whole batched substeps
 * with the resident solves in one cooperative launch and with one launch per
 * iteration at 1/4, 1/2, 1 and 2 times the resident workgroup count, keeps the
 * fastest for stream steps (mode 0 or 1) and the fastest per-iteration width
 * for graph replays, and restores the tanks. */
#define ST_TUNE_MAX 8
typedef struct st_sor_tune {
  uint32_t mode;               /* chosen path (0 resident, 1 per iteration, 2 streaming) */
  uint32_t width;              /* workgroups of the chosen per-iteration launch (0: cooperative) */
  uint32_t n;                  /* candidates timed */
  uint32_t cand_mode[ST_TUNE_MAX];
  uint32_t cand_width[ST_TUNE_MAX];
  float cand_ms[ST_TUNE_MAX];  /* mean time of each candidate (pressure stage, or batched substep) */
} st_sor_tune;
int st_tune_sor(void* sim, const st_params* p, int32_t reps, st_sor_tune* out);
int st_batch_tune_sor(void* batch, const st_params* p, uint32_t np, int32_t reps, st_sor_tune* out);

/* Batched tanks.  No counterpart in the reference, which steps one FlipSim at a
 * time; these advance many independent tanks (each an st_sim of its own size,
 * particles and colliders) with one kernel launch per stage for all of them.
 * Each tank's results are the bits st_step gives it alone.  While in a batch the
 * tanks share the batch's HIP stream (every st_* call on a tank still works and
 * is ordered with the batched steps); st_destroy refuses a tank in a batch. */
/* Group n tanks (at most 256, one device, none in another batch). */
int st_batch_create(void* const* sims, uint32_t n, void** batch);
/* Wait for the batch's work, give each tank its own stream back, free the batch. */
int st_batch_destroy(void* batch);
/* One full substep of every tank (st_step on each), enqueued on the batch's
 * stream.  `p` holds one parameter block for all tanks (np = 1) or one per tank
 * in batch order (np = n).  Tanks with no particles sit the substep out, as
 * st_step does nothing for them.  The narrow band and tile-only steps are not
 * batched (ST_ERR_ARG). */
int st_batch_step(void* batch, const st_params* p, uint32_t np);
/* nsteps batched substeps with the same settings; with use_graph != 0 each is
 * a HIP graph replay, captured the first time its block tables (which fix
 * every launch of the substep) are seen and replayed after that. */
int st_batch_step_n(void* batch, const st_params* p, uint32_t np, int32_t nsteps, int32_t use_graph);
/* Wait for all work of the batch. */
int st_batch_sync(void* batch);
/* Number of tanks. */
int st_batch_size(void* batch, uint32_t* n);
/* The batched pressure solve.  By default (st_batch_step, no graph) every
 * tank's SOR runs resident in one cooperative launch of the device's resident
 * workgroups (one per CU): each tank gets a range of them, enough that its
 * tiles fit in LDS (a tank too big for all of them streams tile windows), in
 * as few phases as the tiles allow, and the same bits as alone.  Graph replays
 * and st_batch_sor_grid(batch, 0) use one launch per iteration instead;
 * st_batch_sor_grid(batch, w) caps the cooperative launch at w workgroups (at
 * least 2, at most the resident count; more phases).  st_batch_sor_path reports whether
 * the last batched substep's solve ran in the one launch. */
int st_batch_sor_grid(void* batch, uint32_t workgroups);
int st_batch_sor_path(void* batch, uint32_t* resident);

/* Snapshots: a copy of the state a substep starts from (particles, tile tables,
 * pool pressure and cell types, time and seed), for re-running substeps from
 * it (the checkpoints of a gradient computation).  No counterpart in the
 * reference.  The pool must not grow between create and restore. */
int st_snapshot_create(void* sim, void** snap);
int st_snapshot_restore(void* sim, void* snap);
int st_snapshot_destroy(void* snap);

/* Adjoint of the substep (the north-star differentiation extension; no
 * counterpart in the reference).  st_adj_record runs one substep exactly as
 * st_step does and keeps what its backward pass reads (a stack of up to
 * max_steps tapes).  st_adj_seed sets dL/d(positions, velocities) of the
 * current state (host float4 arrays in particle-id order, w ignored) and clears
 * the pressure carry; each st_adj_backward then reverses the last recorded
 * substep still on the stack (g2p, projection, the red-black SOR transposed in
 * reverse sweep order, divergence, normFaces, p2g), turning the gradient into
 * that of the state before it; st_adj_grad reads it back.  The derivative is
 * taken at fixed topology (tiles, cell types, counts and branches as the forward
 * pass took them; the fixed-point sums differentiated as the sums they round).
 * Covered: closed boxes without drag or the narrow band, with at most
 * ST_ADJ_MAX_COLS colliders (ST_ERR_ARG otherwise). */
int st_adj_create(void* sim, uint32_t max_steps, void** adj);
int st_adj_destroy(void* adj);
int st_adj_record(void* adj, const st_params* p);
int st_adj_seed(void* adj, const float* g_pos4, const float* g_vel4);
int st_adj_backward(void* adj);
int st_adj_grad(void* adj, float* g_pos4, float* g_vel4);
/* Colliders (at most ST_ADJ_MAX_COLS, rigid records as st_set_colliders takes
 * them; their distance grids must not change between a record and its
 * reversal) are differentiated through the particle push of g2p and the solid
 * faces of normFaces.  After each st_adj_backward, st_adj_col_grad gives, per
 * collider of the substep just reversed, 9 floats: dL/d(centre x, y, z),
 * dL/d(linear velocity), dL/d(angular velocity) of that substep's record
 * (words 20-22, 24-26, 28-30 of the 32; the rotation rows get none).
 * ncols must equal that substep's collider count.  No counterpart in the
 * reference; a keyframed collider's gradient follows from these by the chain
 * rule through its pose curve (shared.js:608-621) and the pose-difference
 * velocity (shared.js:531-537). */
#define ST_ADJ_MAX_COLS 4
int st_adj_col_grad(void* adj, float* g, uint32_t ncols);
/* The transposed solve of st_adj_backward runs as one launch per sweep (the
 * default), or, after st_adj_set_sor_path(adj, 1), as one cooperative launch
 * with a grid barrier between its sweeps where the device allows it (the same
 * sweeps in the same order; (adj, 0) goes back).  On MI355X the one launch is
 * the slower: each barrier has to write back and invalidate the L2 of every
 * XCD, which costs more than a launch boundary.  st_adj_sor_path reports
 * whether the last st_adj_backward ran as one launch.  A barrier time-out (a
 * workgroup that never arrived) is reported as ST_ERR_HIP by the next
 * st_adj_grad or st_adj_col_grad, whose gradients are then not valid; the
 * context falls back to one launch per sweep.  No counterpart in the
 * reference (it has no adjoint). */
int st_adj_set_sor_path(void* adj, uint32_t one_launch);
int st_adj_sor_path(void* adj, uint32_t* one_launch);

#ifdef __cplusplus
}
#endif
#endif /* SPLASHTANK_H */
