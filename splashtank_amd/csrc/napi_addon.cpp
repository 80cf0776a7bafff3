// napi_addon.cpp - Node N-API binding of the C ABI in include/splashtank.h.
//
// Thin by design: each export converts JS numbers, objects and typed arrays to
// the plain pointers and sizes of one st_* call and back.  The JS FlipSim facade
// (splashtank_amd/flip/sim.js) builds the reference's API (src/flip/sim.js) on top.
// Errors from the library are thrown as JS Errors carrying st_last_error().
#include <node_api.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "splashtank.h"

namespace {

napi_value throwErr(napi_env env, const std::string& msg) {
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

#define NAPI_OK(call)                                                   \
  do {                                                                  \
    if ((call) != napi_ok) return throwErr(env, "napi call failed: " #call); \
  } while (0)

#define ST_OK_OR_THROW(call)                                                              \
  do {                                                                                    \
    int rc_ = (call);                                                                     \
    if (rc_ != ST_OK) return throwErr(env, std::string(#call " failed: ") + st_last_error()); \
  } while (0)

bool getArgs(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok) return false;
  for (size_t k = argc; k < want; k++) napi_get_undefined(env, &argv[k]);
  return true;
}

double numProp(napi_env env, napi_value obj, const char* name, double dflt) {
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return dflt;
  napi_value v;
  napi_get_named_property(env, obj, name, &v);
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_boolean) {
    bool b = false;
    napi_get_value_bool(env, v, &b);
    return b ? 1.0 : 0.0;
  }
  if (t != napi_number) return dflt;
  double d = dflt;
  napi_get_value_double(env, v, &d);
  return d;
}

void* handleOf(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok) return nullptr;
  return p;
}

// typed array data pointer and element count
bool typed(napi_env env, napi_value v, void** data, size_t* len, napi_typedarray_type want) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) return false;
  napi_typedarray_type t;
  napi_value ab;
  size_t off;
  if (napi_get_typedarray_info(env, v, &t, len, data, &ab, &off) != napi_ok) return false;
  return t == want;
}

napi_value mkNum(napi_env env, double d) {
  napi_value v;
  napi_create_double(env, d, &v);
  return v;
}

// create({d, ratio, size:[W,H,D], open, maxN, slots, device}) -> handle
napi_value Create(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!getArgs(env, info, 1, argv)) return throwErr(env, "create: bad arguments");
  st_opts o;
  std::memset(&o, 0, sizeof o);
  o.d = numProp(env, argv[0], "d", 0.0);
  o.ratio = numProp(env, argv[0], "ratio", 2.0);
  o.open = uint32_t(numProp(env, argv[0], "open", 0));
  o.max_n = uint32_t(numProp(env, argv[0], "maxN", 0));
  o.slots = uint32_t(numProp(env, argv[0], "slots", 0));
  o.device = int32_t(numProp(env, argv[0], "device", 0));
  bool hasSize = false;
  napi_has_named_property(env, argv[0], "size", &hasSize);
  if (hasSize) {
    napi_value sz;
    napi_get_named_property(env, argv[0], "size", &sz);
    for (uint32_t k = 0; k < 3; k++) {
      napi_value e;
      if (napi_get_element(env, sz, k, &e) == napi_ok) napi_get_value_double(env, e, &o.size[k]);
    }
  }
  void* sim = nullptr;
  ST_OK_OR_THROW(st_create(&o, &sim));
  napi_value h;
  NAPI_OK(napi_create_external(env, sim, nullptr, nullptr, &h));
  return h;
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  getArgs(env, info, 1, argv);
  void* sim = handleOf(env, argv[0]);
  if (sim) ST_OK_OR_THROW(st_destroy(sim));
  return nullptr;
}

napi_value GetDims(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  getArgs(env, info, 1, argv);
  st_dims d;
  ST_OK_OR_THROW(st_get_dims(handleOf(env, argv[0]), &d));
  napi_value o;
  napi_create_object(env, &o);
  const struct { const char* n; double v; } f[] = {
      {"d", d.d}, {"h", d.h}, {"ratio", d.ratio}, {"rest", d.rest}, {"x0", d.x0}, {"y0", d.y0}, {"z0", d.z0},
      {"x1", d.x1}, {"y1", d.y1}, {"z1", d.z1}, {"nx", double(d.nx)}, {"ny", double(d.ny)}, {"nz", double(d.nz)},
      {"tx", double(d.tx)}, {"ty", double(d.ty)}, {"tz", double(d.tz)}, {"ntiles", double(d.ntiles)},
      {"scap", double(d.scap)}, {"scapMax", double(d.scap_max)}, {"maxN", double(d.max_n)}, {"ncell", double(d.ncell)}};
  for (const auto& e : f) napi_set_named_property(env, o, e.n, mkNum(env, e.v));
  return o;
}

// upload(h, first, Float32Array pos4, Float32Array vel4, count)
napi_value Upload(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  getArgs(env, info, 5, argv);
  uint32_t first = 0, count = 0;
  napi_get_value_uint32(env, argv[1], &first);
  void *p = nullptr, *v = nullptr;
  size_t np = 0, nv = 0;
  if (!typed(env, argv[2], &p, &np, napi_float32_array) || !typed(env, argv[3], &v, &nv, napi_float32_array))
    return throwErr(env, "upload: pos and vel must be Float32Arrays");
  napi_get_value_uint32(env, argv[4], &count);
  if (size_t(count) * 4 > np || size_t(count) * 4 > nv) return throwErr(env, "upload: arrays shorter than count*4");
  ST_OK_OR_THROW(st_upload_particles(handleOf(env, argv[0]), first, count, (const float*)p, (const float*)v));
  return nullptr;
}

napi_value SetCount(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  getArgs(env, info, 2, argv);
  uint32_t n = 0;
  napi_get_value_uint32(env, argv[1], &n);
  ST_OK_OR_THROW(st_set_count(handleOf(env, argv[0]), n));
  return nullptr;
}

napi_value Clear(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  getArgs(env, info, 1, argv);
  ST_OK_OR_THROW(st_clear(handleOf(env, argv[0])));
  return nullptr;
}

// setColliders(h, Float32Array cf, nc, Float32Array sdf, sig)
napi_value SetColliders(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  getArgs(env, info, 5, argv);
  void *cf = nullptr, *sdf = nullptr;
  size_t ncf = 0, nsdf = 0;
  if (!typed(env, argv[1], &cf, &ncf, napi_float32_array) || !typed(env, argv[3], &sdf, &nsdf, napi_float32_array))
    return throwErr(env, "setColliders: cf and sdf must be Float32Arrays");
  uint32_t nc = 0;
  napi_get_value_uint32(env, argv[2], &nc);
  if (size_t(nc) * 32 > ncf) return throwErr(env, "setColliders: cf shorter than nc*32");
  double sig = 0;
  napi_get_value_double(env, argv[4], &sig);
  ST_OK_OR_THROW(st_set_colliders(handleOf(env, argv[0]), nc ? (const float*)cf : nullptr, nc, nsdf ? (const float*)sdf : nullptr, nsdf,
                                  uint64_t(sig)));
  return nullptr;
}

// params object -> st_params (field names as in the reference's step(dt, o))
void readParams(napi_env env, napi_value o, st_params* p) {
  st_params_default(p);
  p->dt = float(numProp(env, o, "dt", p->dt));
  p->g = float(numProp(env, o, "g", p->g));
  p->flip = float(numProp(env, o, "flip", p->flip));
  p->comp = float(numProp(env, o, "comp", p->comp));
  p->omega = float(numProp(env, o, "omega", 0));
  p->vmax = float(numProp(env, o, "vmax", 0));
  p->tear = float(numProp(env, o, "tear", 0));
  p->iters = int32_t(numProp(env, o, "iters", p->iters));
  p->drag = float(numProp(env, o, "drag", 0));
  p->foam_tau = float(numProp(env, o, "foamTau", p->foam_tau));
  p->foam_sens = float(numProp(env, o, "foamSens", p->foam_sens));
  p->gust = float(numProp(env, o, "gust", 0));
  p->no_foam = uint32_t(numProp(env, o, "noFoam", 0));
  p->no_cull = uint32_t(numProp(env, o, "noCull", 0));
  p->fast_air = uint32_t(numProp(env, o, "fastAir", 0));
  p->no_solid_cache = uint32_t(numProp(env, o, "noSolidCache", 0));
  p->nb_on = uint32_t(numProp(env, o, "nbOn", 0));
  p->nb_band = float(numProp(env, o, "nbBand", 3));
  p->only_tiles = uint32_t(numProp(env, o, "onlyTiles", 0));
  p->nb_init = uint32_t(numProp(env, o, "nbInit", 0));
  p->nfill = uint32_t(numProp(env, o, "nfill", 0));
  p->nfv = uint32_t(numProp(env, o, "nfv", 0));
  p->sor_mode = uint32_t(numProp(env, o, "sorMode", 0));
  // whitewater and air (o.spr, o.spr.ww, o.air, flattened by the facade)
  p->spr_on = uint32_t(numProp(env, o, "sprOn", 0));
  p->spr_rate = float(numProp(env, o, "sprRate", 0));
  p->spr_min = float(numProp(env, o, "sprMin", 0));
  p->spr_life = float(numProp(env, o, "sprLife", 0));
  p->spr_mist = float(numProp(env, o, "sprMist", 0));
  p->spr_drag = float(numProp(env, o, "sprDrag", 0));
  p->spr_mdrag = float(numProp(env, o, "sprMdrag", 0));
  p->vs = float(numProp(env, o, "vs", 0));
  p->ww_foam_life = float(numProp(env, o, "wwFoamLife", 0));
  p->ww_rise = float(numProp(env, o, "wwRise", 0));
  p->ww_crest = float(numProp(env, o, "wwCrest", p->ww_crest));
  p->ww_impact = float(numProp(env, o, "wwImpact", p->ww_impact));
  p->air_on = uint32_t(numProp(env, o, "airOn", 0));
  p->air_couple = float(numProp(env, o, "airCouple", 0));
  p->air_decay = float(numProp(env, o, "airDecay", 0));
  p->air_blur = float(numProp(env, o, "airBlur", p->air_blur));
  p->drip_on = uint32_t(numProp(env, o, "dripOn", 0));
  p->drip_rate = float(numProp(env, o, "dripRate", 0));
  p->drip_dry = float(numProp(env, o, "dripDry", 0));
  bool has = false;
  napi_has_named_property(env, o, "wind", &has);
  if (has) {
    napi_value w;
    napi_get_named_property(env, o, "wind", &w);
    for (uint32_t k = 0; k < 3; k++) {
      napi_value e;
      double d = 0;
      if (napi_get_element(env, w, k, &e) == napi_ok && napi_get_value_double(env, e, &d) == napi_ok) p->wind[k] = float(d);
    }
  }
  has = false;
  napi_has_named_property(env, o, "fills", &has);
  if (has) {
    napi_value fv;
    napi_get_named_property(env, o, "fills", &fv);
    void* data = nullptr;
    size_t n = 0;
    if (typed(env, fv, &data, &n, napi_float32_array)) std::memcpy(p->fills, data, std::min<size_t>(n, 64) * 4);
  }
  // the ocean (o.ocean, o.time): the facade passes the wave list oceanWaves gives
  p->ocean_on = uint32_t(numProp(env, o, "oceanOn", 0));
  p->ocean_time = float(numProp(env, o, "oceanTime", 0));
  p->ocean_level = float(numProp(env, o, "oceanLevel", 0));
  p->ocean_layer = float(numProp(env, o, "oceanLayer", 0));
  p->rk2 = uint32_t(numProp(env, o, "rk2", 0));
  p->extrap_layers = uint32_t(numProp(env, o, "extrapLayers", 0));
  has = false;
  napi_has_named_property(env, o, "waves", &has);
  if (has) {
    napi_value wv;
    napi_get_named_property(env, o, "waves", &wv);
    void* data = nullptr;
    size_t n = 0;
    if (typed(env, wv, &data, &n, napi_float32_array)) {
      n = std::min<size_t>(n, 8 * ST_OCEAN_MAX_WAVES);
      std::memcpy(p->waves, data, n * 4);
      p->nwave = uint32_t(n / 8);
    }
  }
}

// oceanWaves(oc {height, length, dir, spread, seed}, g, depth) -> Float32Array, 8 floats per wave
napi_value OceanWaves(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  getArgs(env, info, 3, argv);
  st_ocean oc{};
  oc.height = numProp(env, argv[0], "height", 0);
  oc.length = numProp(env, argv[0], "length", 1);
  oc.dir = numProp(env, argv[0], "dir", 0);
  oc.spread = numProp(env, argv[0], "spread", 0);
  oc.seed = int32_t(numProp(env, argv[0], "seed", 0));
  double g = 0, depth = 0;
  napi_get_value_double(env, argv[1], &g);
  napi_get_value_double(env, argv[2], &depth);
  float w[8 * ST_OCEAN_MAX_WAVES];
  uint32_t n = 0;
  ST_OK_OR_THROW(st_ocean_waves(&oc, g, depth, w, ST_OCEAN_MAX_WAVES, &n));
  n = std::min(n, ST_OCEAN_MAX_WAVES);
  napi_value ab, arr;
  void* data = nullptr;
  NAPI_OK(napi_create_arraybuffer(env, size_t(n) * 32, &data, &ab));
  std::memcpy(data, w, size_t(n) * 32);
  NAPI_OK(napi_create_typedarray(env, napi_float32_array, size_t(n) * 8, ab, 0, &arr));
  return arr;
}

napi_value Step(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  getArgs(env, info, 2, argv);
  st_params p;
  readParams(env, argv[1], &p);
  ST_OK_OR_THROW(st_step(handleOf(env, argv[0]), &p));
  return nullptr;
}

napi_value BeginStep(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  getArgs(env, info, 2, argv);
  st_params p;
  readParams(env, argv[1], &p);
  ST_OK_OR_THROW(st_begin_step(handleOf(env, argv[0]), &p));
  return nullptr;
}

napi_value Stage(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  getArgs(env, info, 2, argv);
  int32_t k = 0;
  napi_get_value_int32(env, argv[1], &k);
  ST_OK_OR_THROW(st_stage(handleOf(env, argv[0]), k));
  return nullptr;
}

// readParticles(h, Float32Array dst, n) -> n
napi_value ReadParticles(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  getArgs(env, info, 3, argv);
  void* dst = nullptr;
  size_t len = 0;
  if (!typed(env, argv[1], &dst, &len, napi_float32_array)) return throwErr(env, "readParticles: dst must be a Float32Array");
  uint32_t n = 0;
  napi_get_value_uint32(env, argv[2], &n);
  if (size_t(n) * 4 > len) return throwErr(env, "readParticles: dst shorter than n*4");
  ST_OK_OR_THROW(st_read_particles(handleOf(env, argv[0]), (float*)dst, n));
  return mkNum(env, n);
}

napi_value ReadCtl(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  getArgs(env, info, 1, argv);
  napi_value ab, arr;
  void* data = nullptr;
  NAPI_OK(napi_create_arraybuffer(env, 64, &data, &ab));
  ST_OK_OR_THROW(st_read_ctl(handleOf(env, argv[0]), (uint32_t*)data));
  NAPI_OK(napi_create_typedarray(env, napi_uint32_array, 16, ab, 0, &arr));
  return arr;
}

napi_value Grow(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  getArgs(env, info, 2, argv);
  uint32_t n = 0;
  napi_get_value_uint32(env, argv[1], &n);
  ST_OK_OR_THROW(st_grow(handleOf(env, argv[0]), n));
  return nullptr;
}

// dump(h, what) -> ArrayBuffer (canonical layout, see splashtank.h)
napi_value Dump(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  getArgs(env, info, 2, argv);
  int32_t what = 0;
  napi_get_value_int32(env, argv[1], &what);
  void* sim = handleOf(env, argv[0]);
  size_t bytes = 0;
  ST_OK_OR_THROW(st_dump_size(sim, what, &bytes));
  napi_value ab;
  void* data = nullptr;
  NAPI_OK(napi_create_arraybuffer(env, bytes ? bytes : 4, &data, &ab));
  if (bytes) ST_OK_OR_THROW(st_dump(sim, what, data, bytes));
  return ab;
}

napi_value Sync(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  getArgs(env, info, 1, argv);
  ST_OK_OR_THROW(st_sync(handleOf(env, argv[0])));
  return nullptr;
}

napi_value MemBytes(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  getArgs(env, info, 1, argv);
  uint64_t b = 0;
  ST_OK_OR_THROW(st_mem_bytes(handleOf(env, argv[0]), &b));
  return mkNum(env, double(b));
}

// buildSdf(Float64Array v (3 per vertex), Uint32Array tris (3 per triangle), cell, margin)
//   -> {dims: [nx, ny, nz], origin: [x, y, z], cell, data: Float32Array}
napi_value BuildSdf(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  getArgs(env, info, 4, argv);
  void *v = nullptr, *t = nullptr;
  size_t nv = 0, nt = 0;
  if (!typed(env, argv[0], &v, &nv, napi_float64_array) || !typed(env, argv[1], &t, &nt, napi_uint32_array))
    return throwErr(env, "buildSdf: vertices must be a Float64Array and triangles a Uint32Array");
  double cell = 0.025, margin = 0.16;
  napi_get_value_double(env, argv[2], &cell);
  napi_get_value_double(env, argv[3], &margin);
  st_sdf_grid g;
  if (st_build_sdf((const double*)v, uint32_t(nv / 3), (const uint32_t*)t, uint32_t(nt / 3), cell, margin, &g, nullptr, 0) != ST_OK)
    return throwErr(env, "buildSdf: bad mesh");
  const size_t n = size_t(g.dims[0]) * g.dims[1] * g.dims[2];
  void* data = nullptr;
  napi_value ab, ta, out, dims, org;
  napi_create_arraybuffer(env, n * 4, &data, &ab);
  if (st_build_sdf((const double*)v, uint32_t(nv / 3), (const uint32_t*)t, uint32_t(nt / 3), cell, margin, &g, (float*)data, n) != ST_OK)
    return throwErr(env, "buildSdf: failed");
  napi_create_typedarray(env, napi_float32_array, n, ab, 0, &ta);
  napi_create_object(env, &out);
  napi_create_array_with_length(env, 3, &dims);
  napi_create_array_with_length(env, 3, &org);
  for (uint32_t k = 0; k < 3; k++) {
    napi_set_element(env, dims, k, mkNum(env, double(g.dims[k])));
    napi_set_element(env, org, k, mkNum(env, g.origin[k]));
  }
  napi_set_named_property(env, out, "dims", dims);
  napi_set_named_property(env, out, "origin", org);
  napi_set_named_property(env, out, "cell", mkNum(env, g.cell));
  napi_set_named_property(env, out, "data", ta);
  return out;
}

// packView(h, dispMax) -> {layout: {nW, nS, nD, words, oWp, ...}, words: Uint32Array} (st_pack_view)
napi_value PackView(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  getArgs(env, info, 2, argv);
  uint32_t dispMax = 0;
  napi_get_value_uint32(env, argv[1], &dispMax);
  void* h = handleOf(env, argv[0]);
  st_view_layout L;
  ST_OK_OR_THROW(st_pack_view(h, dispMax, nullptr, 0, &L));
  void* data = nullptr;
  napi_value ab, ta, out, lay;
  napi_create_arraybuffer(env, size_t(std::max(L.words, 1u)) * 4, &data, &ab);
  ST_OK_OR_THROW(st_pack_view(h, dispMax, (uint32_t*)data, L.words, &L));
  napi_create_typedarray(env, napi_uint32_array, L.words, ab, 0, &ta);
  napi_create_object(env, &lay);
  const struct { const char* k; uint32_t v; } f[] = {{"nW", L.nW}, {"nS", L.nS}, {"nD", L.nD}, {"words", L.words}, {"oWp", L.oWp},
                                                      {"oWs", L.oWs}, {"oWf", L.oWf}, {"oSp", L.oSp}, {"oSk", L.oSk}, {"oDp", L.oDp},
                                                      {"oDs", L.oDs}, {"stride", L.stride}, {"fromPos", L.from_pos}};
  for (const auto& e : f) napi_set_named_property(env, lay, e.k, mkNum(env, double(e.v)));
  napi_create_object(env, &out);
  napi_set_named_property(env, out, "layout", lay);
  napi_set_named_property(env, out, "words", ta);
  return out;
}

// setDrips(h, Float32Array samples (8 per sample))
napi_value SetDrips(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  getArgs(env, info, 2, argv);
  void* d = nullptr;
  size_t n = 0;
  if (!typed(env, argv[1], &d, &n, napi_float32_array)) return throwErr(env, "setDrips: samples must be a Float32Array");
  ST_OK_OR_THROW(st_set_drips(handleOf(env, argv[0]), (const float*)d, uint32_t(n / 8)));
  return nullptr;
}

// setSprayCap(h, n)
napi_value SetSprayCap(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  getArgs(env, info, 2, argv);
  uint32_t n = 0;
  napi_get_value_uint32(env, argv[1], &n);
  ST_OK_OR_THROW(st_set_spray_cap(handleOf(env, argv[0]), n));
  return nullptr;
}

// ---- asynchronous reads (st_read_begin / st_read_wait / st_read_free) ----
// readAsync(h, what, dst, n) -> Promise.  The copy is enqueued on the sim's
// stream at once (behind every step enqueued before); a libuv worker thread
// waits for it and copies it out (napi_async_work), so the event loop runs
// meanwhile; the Promise settles on the main thread.  what 0: particles into
// dst (a Float32Array of >= 4n floats, kept alive until then), resolves n;
// 1: the 16 control words, resolves a Uint32Array; 2: whitewater slots,
// resolves {n, p, v} or null; 3: deep water points, resolves {n, p} or null.
struct ReadWork {
  napi_async_work work = nullptr;
  napi_deferred deferred = nullptr;
  napi_ref dstRef = nullptr;
  void* ticket = nullptr;
  uint32_t what = 0, room = 0, n = 0;
  float* dst = nullptr;
  std::vector<float> a, b;
  uint32_t ctl[16] = {};
  int rc = 0;
  std::string err;
};

void readExecute(napi_env, void* data) {
  ReadWork* w = (ReadWork*)data;
  void *o0 = nullptr, *o1 = nullptr;
  if (w->what == ST_READ_PARTICLES) {
    o0 = w->dst;
  } else if (w->what == ST_READ_CTL) {
    o0 = w->ctl;
    w->room = 16;
  } else if (w->what == ST_READ_SPRAY) {
    w->a.resize(size_t(std::max(w->room, 1u)) * 4);
    w->b.resize(size_t(std::max(w->room, 1u)) * 4);
    o0 = w->a.data();
    o1 = w->b.data();
  } else {
    w->room = ST_DEEP_MAX;
    w->a.resize(size_t(ST_DEEP_MAX) * 4);
    o0 = w->a.data();
  }
  w->rc = st_read_wait(w->ticket, o0, o1, w->room, &w->n);
  if (w->rc != ST_OK) w->err = st_last_error();  // this thread's error text
}

napi_value floats(napi_env env, const float* src, uint32_t n4) {
  void* data = nullptr;
  napi_value ab, ta;
  napi_create_arraybuffer(env, size_t(n4) * 4, &data, &ab);
  if (n4) std::memcpy(data, src, size_t(n4) * 4);
  napi_create_typedarray(env, napi_float32_array, n4, ab, 0, &ta);
  return ta;
}

void readComplete(napi_env env, napi_status status, void* data) {
  ReadWork* w = (ReadWork*)data;
  st_read_free(w->ticket);
  if (w->dstRef) napi_delete_reference(env, w->dstRef);
  napi_value out;
  if (status != napi_ok || w->rc != ST_OK) {
    napi_value msg;
    const std::string m = "readAsync failed: " + (w->err.empty() ? std::string("cancelled") : w->err);
    napi_create_string_utf8(env, m.c_str(), NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, nullptr, msg, &out);
    napi_reject_deferred(env, w->deferred, out);
  } else {
    if (w->what == ST_READ_PARTICLES) {
      out = mkNum(env, double(w->n));
    } else if (w->what == ST_READ_CTL) {
      void* d = nullptr;
      napi_value ab;
      napi_create_arraybuffer(env, 64, &d, &ab);
      std::memcpy(d, w->ctl, 64);
      napi_create_typedarray(env, napi_uint32_array, 16, ab, 0, &out);
    } else if (!w->n) {
      napi_get_null(env, &out);
    } else {
      napi_create_object(env, &out);
      napi_set_named_property(env, out, "n", mkNum(env, double(w->n)));
      napi_set_named_property(env, out, "p", floats(env, w->a.data(), w->n * 4));
      if (w->what == ST_READ_SPRAY) napi_set_named_property(env, out, "v", floats(env, w->b.data(), w->n * 4));
    }
    napi_resolve_deferred(env, w->deferred, out);
  }
  napi_delete_async_work(env, w->work);
  delete w;
}

napi_value ReadAsync(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  getArgs(env, info, 4, argv);
  void* h = handleOf(env, argv[0]);
  uint32_t what = 0, n = 0;
  napi_get_value_uint32(env, argv[1], &what);
  napi_get_value_uint32(env, argv[3], &n);
  if (what > ST_READ_DEEP) return throwErr(env, "readAsync: unknown result");
  std::unique_ptr<ReadWork> w(new ReadWork());
  w->what = what;
  w->room = n;
  if (what == ST_READ_PARTICLES) {
    void* dst = nullptr;
    size_t len = 0;
    if (!typed(env, argv[2], &dst, &len, napi_float32_array)) return throwErr(env, "readAsync: dst must be a Float32Array");
    if (size_t(n) * 4 > len) return throwErr(env, "readAsync: dst shorter than n*4");
    w->dst = (float*)dst;
    NAPI_OK(napi_create_reference(env, argv[2], 1, &w->dstRef));
  }
  int rc = st_read_begin(h, what, n, &w->ticket);
  if (rc != ST_OK) {
    if (w->dstRef) napi_delete_reference(env, w->dstRef);
    return throwErr(env, std::string("st_read_begin failed: ") + st_last_error());
  }
  napi_value promise, name;
  NAPI_OK(napi_create_promise(env, &w->deferred, &promise));
  NAPI_OK(napi_create_string_utf8(env, "splashtank.readAsync", NAPI_AUTO_LENGTH, &name));
  NAPI_OK(napi_create_async_work(env, nullptr, name, readExecute, readComplete, w.get(), &w->work));
  NAPI_OK(napi_queue_async_work(env, w->work));
  w.release();
  return promise;
}

// readSpray(h, room) -> {n, p: Float32Array(4n), v: Float32Array(4n)} or null without a pool
napi_value ReadDeep(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  getArgs(env, info, 1, argv);
  std::vector<float> p(size_t(ST_DEEP_MAX) * 4);
  uint32_t n = 0;
  ST_OK_OR_THROW(st_read_deep(handleOf(env, argv[0]), p.data(), ST_DEEP_MAX, &n));
  napi_value out;
  if (!n) {
    napi_get_null(env, &out);
    return out;
  }
  napi_create_object(env, &out);
  napi_set_named_property(env, out, "n", mkNum(env, double(n)));
  void* data = nullptr;
  napi_value ab, ta;
  napi_create_arraybuffer(env, size_t(n) * 16, &data, &ab);
  std::memcpy(data, p.data(), size_t(n) * 16);
  napi_create_typedarray(env, napi_float32_array, size_t(n) * 4, ab, 0, &ta);
  napi_set_named_property(env, out, "p", ta);
  return out;
}
napi_value ReadSpray(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  getArgs(env, info, 2, argv);
  uint32_t room = 0;
  napi_get_value_uint32(env, argv[1], &room);
  std::vector<float> p(size_t(std::max(room, 1u)) * 4), v(size_t(std::max(room, 1u)) * 4);
  uint32_t n = 0;
  ST_OK_OR_THROW(st_read_spray(handleOf(env, argv[0]), p.data(), v.data(), room, &n));
  napi_value out;
  if (!n) {
    napi_get_null(env, &out);
    return out;
  }
  napi_create_object(env, &out);
  napi_set_named_property(env, out, "n", mkNum(env, double(n)));
  const char* names[2] = {"p", "v"};
  float* src[2] = {p.data(), v.data()};
  for (int q = 0; q < 2; q++) {
    void* data = nullptr;
    napi_value ab, ta;
    napi_create_arraybuffer(env, size_t(n) * 16, &data, &ab);
    std::memcpy(data, src[q], size_t(n) * 16);
    napi_create_typedarray(env, napi_float32_array, size_t(n) * 4, ab, 0, &ta);
    napi_set_named_property(env, out, names[q], ta);
  }
  return out;
}

// solveNorms(h) -> {resL2, resMax, rhsL2, divL2, divMax, errL2, errMax, fluid} (st_solve_norms)
napi_value SolveNorms(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  getArgs(env, info, 1, argv);
  st_norms n;
  ST_OK_OR_THROW(st_solve_norms(handleOf(env, argv[0]), &n));
  napi_value o;
  napi_create_object(env, &o);
  const struct { const char* k; double v; } f[] = {{"resL2", n.res_l2}, {"resMax", n.res_max}, {"rhsL2", n.rhs_l2},
                                                    {"divL2", n.div_l2}, {"divMax", n.div_max}, {"errL2", n.err_l2},
                                                    {"errMax", n.err_max}, {"fluid", n.fluid}};
  for (const auto& e : f) napi_set_named_property(env, o, e.k, mkNum(env, e.v));
  return o;
}

napi_value Version(napi_env env, napi_callback_info) {
  napi_value v;
  napi_create_string_utf8(env, st_version(), NAPI_AUTO_LENGTH, &v);
  return v;
}

// batchCreate([handles]) -> batch handle (st_batch_create)
napi_value BatchCreate(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  getArgs(env, info, 1, argv);
  bool isArr = false;
  napi_is_array(env, argv[0], &isArr);
  if (!isArr) return throwErr(env, "batchCreate: expects an array of sim handles");
  uint32_t n = 0;
  napi_get_array_length(env, argv[0], &n);
  std::vector<void*> sims(n);
  for (uint32_t k = 0; k < n; k++) {
    napi_value e;
    NAPI_OK(napi_get_element(env, argv[0], k, &e));
    sims[k] = handleOf(env, e);
  }
  void* b = nullptr;
  ST_OK_OR_THROW(st_batch_create(sims.data(), n, &b));
  napi_value h;
  NAPI_OK(napi_create_external(env, b, nullptr, nullptr, &h));
  return h;
}

napi_value BatchDestroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  getArgs(env, info, 1, argv);
  void* b = handleOf(env, argv[0]);
  if (b) ST_OK_OR_THROW(st_batch_destroy(b));
  return nullptr;
}

// batchStep(batch, [params per tank]) (st_batch_step)
napi_value BatchStep(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  getArgs(env, info, 2, argv);
  uint32_t n = 0;
  napi_get_array_length(env, argv[1], &n);
  std::vector<st_params> p(n);
  for (uint32_t k = 0; k < n; k++) {
    napi_value e;
    NAPI_OK(napi_get_element(env, argv[1], k, &e));
    readParams(env, e, &p[k]);
  }
  ST_OK_OR_THROW(st_batch_step(handleOf(env, argv[0]), p.data(), n));
  return nullptr;
}

napi_value BatchSync(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  getArgs(env, info, 1, argv);
  ST_OK_OR_THROW(st_batch_sync(handleOf(env, argv[0])));
  return nullptr;
}

// ---- the extensions: graph replay, the SOR autotuner, snapshots and the adjoint ----

int32_t intArg(napi_env env, napi_value v, int32_t dflt) {
  napi_valuetype t;
  if (napi_typeof(env, v, &t) != napi_ok || t != napi_number) return dflt;
  int32_t x = dflt;
  napi_get_value_int32(env, v, &x);
  return x;
}

// stepN(h, params, nsteps, useGraph) (st_step_n)
napi_value StepN(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  getArgs(env, info, 4, argv);
  st_params p;
  readParams(env, argv[1], &p);
  ST_OK_OR_THROW(st_step_n(handleOf(env, argv[0]), &p, intArg(env, argv[2], 1), intArg(env, argv[3], 0)));
  return nullptr;
}

// batchStepN(batch, [params per tank], nsteps, useGraph) (st_batch_step_n)
napi_value BatchStepN(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  getArgs(env, info, 4, argv);
  uint32_t n = 0;
  napi_get_array_length(env, argv[1], &n);
  std::vector<st_params> p(n);
  for (uint32_t k = 0; k < n; k++) {
    napi_value e;
    NAPI_OK(napi_get_element(env, argv[1], k, &e));
    readParams(env, e, &p[k]);
  }
  ST_OK_OR_THROW(st_batch_step_n(handleOf(env, argv[0]), p.data(), n, intArg(env, argv[2], 1), intArg(env, argv[3], 0)));
  return nullptr;
}

napi_value tuneResult(napi_env env, const st_sor_tune& t) {
  napi_value o, arr;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "mode", mkNum(env, t.mode));
  napi_set_named_property(env, o, "width", mkNum(env, t.width));
  napi_create_array_with_length(env, t.n, &arr);
  for (uint32_t k = 0; k < t.n && k < ST_TUNE_MAX; k++) {
    napi_value c;
    napi_create_object(env, &c);
    napi_set_named_property(env, c, "mode", mkNum(env, t.cand_mode[k]));
    napi_set_named_property(env, c, "width", mkNum(env, t.cand_width[k]));
    napi_set_named_property(env, c, "ms", mkNum(env, t.cand_ms[k]));
    napi_set_element(env, arr, k, c);
  }
  napi_set_named_property(env, o, "candidates", arr);
  return o;
}

// tuneSor(h, params, reps) -> {mode, width, candidates: [{mode, width, ms}]} (st_tune_sor)
napi_value TuneSor(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  getArgs(env, info, 3, argv);
  st_params p;
  readParams(env, argv[1], &p);
  st_sor_tune t{};
  ST_OK_OR_THROW(st_tune_sor(handleOf(env, argv[0]), &p, intArg(env, argv[2], 3), &t));
  return tuneResult(env, t);
}

// batchTuneSor(batch, [params per tank], reps) (st_batch_tune_sor)
napi_value BatchTuneSor(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  getArgs(env, info, 3, argv);
  uint32_t n = 0;
  napi_get_array_length(env, argv[1], &n);
  std::vector<st_params> p(n);
  for (uint32_t k = 0; k < n; k++) {
    napi_value e;
    NAPI_OK(napi_get_element(env, argv[1], k, &e));
    readParams(env, e, &p[k]);
  }
  st_sor_tune t{};
  ST_OK_OR_THROW(st_batch_tune_sor(handleOf(env, argv[0]), p.data(), n, intArg(env, argv[2], 3), &t));
  return tuneResult(env, t);
}

napi_value external(napi_env env, void* p) {
  napi_value h;
  if (napi_create_external(env, p, nullptr, nullptr, &h) != napi_ok) return throwErr(env, "napi_create_external failed");
  return h;
}

// snapshotCreate(h) -> snapshot; snapshotRestore(h, snapshot); snapshotDestroy(snapshot)
napi_value SnapshotCreate(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  getArgs(env, info, 1, argv);
  void* snap = nullptr;
  ST_OK_OR_THROW(st_snapshot_create(handleOf(env, argv[0]), &snap));
  return external(env, snap);
}
napi_value SnapshotRestore(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  getArgs(env, info, 2, argv);
  ST_OK_OR_THROW(st_snapshot_restore(handleOf(env, argv[0]), handleOf(env, argv[1])));
  return nullptr;
}
napi_value SnapshotDestroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  getArgs(env, info, 1, argv);
  void* snap = handleOf(env, argv[0]);
  if (snap) ST_OK_OR_THROW(st_snapshot_destroy(snap));
  return nullptr;
}

// adjCreate(h, maxSteps) -> adjoint; adjRecord(adj, params); adjSeed(adj, gPos4, gVel4);
// adjBackward(adj); adjGrad(adj, gPos4, gVel4) fills them; adjColGrad(adj, ncols) -> Float32Array(9 ncols);
// adjDestroy(adj)  (st_adj_*)
napi_value AdjCreate(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  getArgs(env, info, 2, argv);
  void* adj = nullptr;
  ST_OK_OR_THROW(st_adj_create(handleOf(env, argv[0]), uint32_t(intArg(env, argv[1], 1)), &adj));
  return external(env, adj);
}
napi_value AdjRecord(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  getArgs(env, info, 2, argv);
  st_params p;
  readParams(env, argv[1], &p);
  ST_OK_OR_THROW(st_adj_record(handleOf(env, argv[0]), &p));
  return nullptr;
}
napi_value AdjSeed(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  getArgs(env, info, 3, argv);
  void *gp = nullptr, *gv = nullptr;
  size_t np = 0, nv = 0;
  if (!typed(env, argv[1], &gp, &np, napi_float32_array) || !typed(env, argv[2], &gv, &nv, napi_float32_array) || np != nv)
    return throwErr(env, "adjSeed: expects two Float32Arrays of the same length (4 per particle)");
  ST_OK_OR_THROW(st_adj_seed(handleOf(env, argv[0]), (const float*)gp, (const float*)gv));
  return nullptr;
}
napi_value AdjBackward(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  getArgs(env, info, 1, argv);
  ST_OK_OR_THROW(st_adj_backward(handleOf(env, argv[0])));
  return nullptr;
}
napi_value AdjGrad(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  getArgs(env, info, 3, argv);
  void *gp = nullptr, *gv = nullptr;
  size_t np = 0, nv = 0;
  if (!typed(env, argv[1], &gp, &np, napi_float32_array) || !typed(env, argv[2], &gv, &nv, napi_float32_array) || np != nv)
    return throwErr(env, "adjGrad: expects two Float32Arrays of the same length (4 per particle)");
  ST_OK_OR_THROW(st_adj_grad(handleOf(env, argv[0]), (float*)gp, (float*)gv));
  return nullptr;
}
napi_value AdjColGrad(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  getArgs(env, info, 2, argv);
  const uint32_t nc = uint32_t(intArg(env, argv[1], 0));
  std::vector<float> g(size_t(nc) * 9 + 1);
  ST_OK_OR_THROW(st_adj_col_grad(handleOf(env, argv[0]), g.data(), nc));
  napi_value ab, arr;
  void* data = nullptr;
  NAPI_OK(napi_create_arraybuffer(env, size_t(nc) * 36, &data, &ab));
  if (nc) std::memcpy(data, g.data(), size_t(nc) * 36);
  NAPI_OK(napi_create_typedarray(env, napi_float32_array, size_t(nc) * 9, ab, 0, &arr));
  return arr;
}
napi_value AdjDestroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  getArgs(env, info, 1, argv);
  void* adj = handleOf(env, argv[0]);
  if (adj) ST_OK_OR_THROW(st_adj_destroy(adj));
  return nullptr;
}

napi_value Init(napi_env env, napi_value exports) {
  const struct { const char* n; napi_callback f; } fns[] = {
      {"create", Create}, {"destroy", Destroy}, {"getDims", GetDims}, {"upload", Upload}, {"setCount", SetCount},
      {"clear", Clear}, {"setColliders", SetColliders}, {"step", Step}, {"beginStep", BeginStep}, {"stage", Stage},
      {"readParticles", ReadParticles}, {"readCtl", ReadCtl}, {"grow", Grow}, {"dump", Dump}, {"sync", Sync},
      {"memBytes", MemBytes}, {"solveNorms", SolveNorms}, {"setSprayCap", SetSprayCap}, {"buildSdf", BuildSdf}, {"oceanWaves", OceanWaves}, {"packView", PackView}, {"setDrips", SetDrips}, {"readSpray", ReadSpray}, {"readDeep", ReadDeep}, {"readAsync", ReadAsync}, {"version", Version}, {"batchCreate", BatchCreate}, {"batchDestroy", BatchDestroy},
      {"batchStep", BatchStep}, {"batchSync", BatchSync}, {"stepN", StepN}, {"batchStepN", BatchStepN}, {"tuneSor", TuneSor},
      {"batchTuneSor", BatchTuneSor}, {"snapshotCreate", SnapshotCreate}, {"snapshotRestore", SnapshotRestore},
      {"snapshotDestroy", SnapshotDestroy}, {"adjCreate", AdjCreate}, {"adjRecord", AdjRecord}, {"adjSeed", AdjSeed},
      {"adjBackward", AdjBackward}, {"adjGrad", AdjGrad}, {"adjColGrad", AdjColGrad}, {"adjDestroy", AdjDestroy}};
  for (const auto& e : fns) {
    napi_value f;
    napi_create_function(env, e.n, NAPI_AUTO_LENGTH, e.f, nullptr, &f);
    napi_set_named_property(env, exports, e.n, f);
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
