// nozzle_napi.cpp — Node-API addon binding the C ABI (include/nozzle.h) for the
// JavaScript drop-in classes in gpu_solver.js. Every binding is a thin argument
// conversion; a non-zero status becomes a thrown JS Error carrying nz_last_error(),
// as the reference throws from GPUSolver (solver.js:680, 713, 723, 748).
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "nozzle.h"

namespace {

#define NAPI_OK(call)                                   \
  do {                                                  \
    if ((call) != napi_ok) {                            \
      napi_throw_error(env, nullptr, "N-API failure");  \
      return nullptr;                                   \
    }                                                   \
  } while (0)

napi_value throw_rc(napi_env env, int rc, const char* what) {
  std::string msg = std::string(what) + ": " + nz_last_error() + " (code " + std::to_string(rc) + ")";
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

struct Args {
  size_t n = 16;
  napi_value v[16];
  napi_value self;
};
bool get_args(napi_env env, napi_callback_info info, Args& a) {
  return napi_get_cb_info(env, info, &a.n, a.v, &a.self, nullptr) == napi_ok;
}

nz_solver* solver_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok) return nullptr;
  return static_cast<nz_solver*>(p);
}
nz_setup* setup_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok) return nullptr;
  return static_cast<nz_setup*>(p);
}
double num(napi_env env, napi_value v, double dflt = 0) {
  double d = dflt;
  napi_valuetype t;
  if (napi_typeof(env, v, &t) != napi_ok) return dflt;
  if (t == napi_number) napi_get_value_double(env, v, &d);
  else if (t == napi_boolean) { bool b = false; napi_get_value_bool(env, v, &b); d = b ? 1 : 0; }
  return d;
}
napi_value prop(napi_env env, napi_value obj, const char* name) {
  napi_value v = nullptr;
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return nullptr;
  napi_get_named_property(env, obj, name, &v);
  return v;
}
double pnum(napi_env env, napi_value obj, const char* name, double dflt = 0) {
  napi_value v = prop(env, obj, name);
  return v ? num(env, v, dflt) : dflt;
}
void parr(napi_env env, napi_value obj, const char* name, float* out, int n) {
  napi_value v = prop(env, obj, name);
  if (!v) return;
  for (int k = 0; k < n; ++k) {
    napi_value e;
    if (napi_get_element(env, v, k, &e) == napi_ok) out[k] = (float)num(env, e);
  }
}
// a Float32Array / Uint8Array argument's data
template <class T>
T* typed_data(napi_env env, napi_value v, size_t* len) {
  napi_typedarray_type type;
  size_t length = 0, offset = 0;
  void* data = nullptr;
  napi_value buf;
  if (napi_get_typedarray_info(env, v, &type, &length, &data, &buf, &offset) != napi_ok) return nullptr;
  if (len) *len = length;
  return static_cast<T*>(data);
}
napi_value new_f32(napi_env env, const float* src, size_t n) {
  void* data = nullptr;
  napi_value ab, ta;
  napi_create_arraybuffer(env, n * 4, &data, &ab);
  if (src && n) std::memcpy(data, src, n * 4);
  napi_create_typedarray(env, napi_float32_array, n, ab, 0, &ta);
  return ta;
}
napi_value new_u8(napi_env env, const uint8_t* src, size_t n) {
  void* data = nullptr;
  napi_value ab, ta;
  napi_create_arraybuffer(env, n, &data, &ab);
  if (src && n) std::memcpy(data, src, n);
  napi_create_typedarray(env, napi_uint8_array, n, ab, 0, &ta);
  return ta;
}
napi_value mk_num(napi_env env, double d) {
  napi_value v;
  napi_create_double(env, d, &v);
  return v;
}
napi_value mk_bool(napi_env env, bool b) {
  napi_value v;
  napi_get_boolean(env, b, &v);
  return v;
}
napi_value mk_null(napi_env env) {
  napi_value v;
  napi_get_null(env, &v);
  return v;
}
void set(napi_env env, napi_value obj, const char* name, napi_value v) { napi_set_named_property(env, obj, name, v); }
napi_value f32_array_js(napi_env env, const float* a, int n) {
  napi_value arr;
  napi_create_array_with_length(env, n, &arr);
  for (int k = 0; k < n; ++k) napi_set_element(env, arr, k, mk_num(env, a[k]));
  return arr;
}

// the reference's P object (simsetup.js:274-310) -> nz_params
void params_from_js(napi_env env, napi_value P, nz_params& p) {
  std::memset(&p, 0, sizeof(p));
  parr(env, P, "amb", p.amb, 4);
  parr(env, P, "res", p.res, 3);
  parr(env, P, "inlet", p.inlet, 4);
  p.pStart = (float)pnum(env, P, "pStart");
  p.ramp = (float)pnum(env, P, "ramp");
  p.RcCells = (float)pnum(env, P, "RcCells");
  p.inj = pnum(env, P, "inj") != 0;
  p.order2 = pnum(env, P, "order2", 1) != 0;
  p.hllBlend = (float)pnum(env, P, "hllBlend", 1);
  parr(env, P, "ambCons", p.ambCons, 4);
  parr(env, P, "ambB", p.ambB, 4);
  parr(env, P, "gamR", p.gamR, 2);
  p.uInlet0 = (float)pnum(env, P, "uInlet0");
  p.vMax = (float)pnum(env, P, "vMax");
  p.cfl = (float)pnum(env, P, "cfl", 0.6);
  p.dtMax = (float)pnum(env, P, "dtMax", 10);
  p.wFloor = (float)pnum(env, P, "wFloor");
  p.visc = pnum(env, P, "visc") != 0;
  p.chem = pnum(env, P, "chem") != 0;
  parr(env, P, "gas", p.gas, 4);
  parr(env, P, "Rs", p.Rs, 3);
  parr(env, P, "chamberNu", p.chamberNu, 2);
  parr(env, P, "cap", p.cap, 2);
  parr(env, P, "chemK", p.chemK, 4);
  p.Tw = (float)pnum(env, P, "Tw");
  napi_value sp = prop(env, P, "spray");
  napi_valuetype t = napi_undefined;
  if (sp) napi_typeof(env, sp, &t);
  if (sp && t == napi_object) {
    p.sprayMhat = (float)pnum(env, sp, "mhat");
    p.sprayTScale = (float)pnum(env, sp, "tScale");
  }
  napi_value d = prop(env, P, "injData");
  t = napi_undefined;
  if (d) napi_typeof(env, d, &t);
  if (d && t == napi_object) {
    parr(env, d, "fuelVec", p.fuelVec, 4);
    parr(env, d, "oxVec", p.oxVec, 4);
    p.eF = (float)pnum(env, d, "eF");
    p.eO = (float)pnum(env, d, "eO");
    p.Ge = (float)pnum(env, d, "Ge");
  }
}
napi_value params_to_js(napi_env env, const nz_params& p, const double* d64) {
  (void)d64;
  napi_value P;
  napi_create_object(env, &P);
  set(env, P, "amb", f32_array_js(env, p.amb, 4));
  set(env, P, "res", f32_array_js(env, p.res, 3));
  set(env, P, "inlet", f32_array_js(env, p.inlet, 4));
  set(env, P, "pStart", mk_num(env, p.pStart));
  set(env, P, "ramp", mk_num(env, p.ramp));
  set(env, P, "RcCells", mk_num(env, p.RcCells));
  set(env, P, "inj", mk_bool(env, p.inj != 0));
  set(env, P, "order2", mk_bool(env, p.order2 != 0));
  set(env, P, "hllBlend", mk_num(env, p.hllBlend));
  set(env, P, "ambCons", f32_array_js(env, p.ambCons, 4));
  set(env, P, "ambB", f32_array_js(env, p.ambB, 4));
  set(env, P, "gamR", f32_array_js(env, p.gamR, 2));
  set(env, P, "uInlet0", mk_num(env, p.uInlet0));
  set(env, P, "vMax", mk_num(env, p.vMax));
  set(env, P, "cfl", mk_num(env, p.cfl));
  set(env, P, "dtMax", mk_num(env, p.dtMax));
  set(env, P, "wFloor", mk_num(env, p.wFloor));
  set(env, P, "visc", mk_bool(env, p.visc != 0));
  set(env, P, "chem", mk_bool(env, p.chem != 0));
  set(env, P, "gas", f32_array_js(env, p.gas, 4));
  set(env, P, "Rs", f32_array_js(env, p.Rs, 3));
  set(env, P, "chamberNu", f32_array_js(env, p.chamberNu, 2));
  set(env, P, "cap", f32_array_js(env, p.cap, 2));
  set(env, P, "chemK", f32_array_js(env, p.chemK, 4));
  set(env, P, "Tw", mk_num(env, p.Tw));
  if (p.inj) {
    napi_value sp, d;
    napi_create_object(env, &sp);
    set(env, sp, "mhat", mk_num(env, p.sprayMhat));
    set(env, sp, "tScale", mk_num(env, p.sprayTScale));
    set(env, P, "spray", sp);
    napi_create_object(env, &d);
    set(env, d, "fuelVec", f32_array_js(env, p.fuelVec, 4));
    set(env, d, "oxVec", f32_array_js(env, p.oxVec, 4));
    set(env, d, "eF", mk_num(env, p.eF));
    set(env, d, "eO", mk_num(env, p.eO));
    set(env, d, "Ge", mk_num(env, p.Ge));
    set(env, P, "injData", d);
  } else {
    set(env, P, "spray", mk_null(env));
    set(env, P, "injData", mk_null(env));
  }
  return P;
}

// ------------------------------------------------------------------ solver
napi_value Create(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a)) return nullptr;
  const int prec = a.n > 0 ? (int)num(env, a.v[0], 32) : 32;
  nz_solver* s = nullptr;
  if (int rc = nz_create_ex(&s, prec)) return throw_rc(env, rc, "GPUSolver");
  napi_value ext;
  NAPI_OK(napi_create_external(env, s, nullptr, nullptr, &ext));
  return ext;
}
napi_value Destroy(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 1) return nullptr;
  nz_destroy(solver_of(env, a.v[0]));
  return nullptr;
}
napi_value Resize(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 3) return throw_rc(env, -1, "resize");
  if (int rc = nz_resize(solver_of(env, a.v[0]), (int)num(env, a.v[1]), (int)num(env, a.v[2]))) return throw_rc(env, rc, "resize");
  return nullptr;
}
napi_value SetMask(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 2) return throw_rc(env, -1, "setMask");
  uint8_t* m = typed_data<uint8_t>(env, a.v[1], nullptr);
  if (int rc = nz_set_mask(solver_of(env, a.v[0]), m)) return throw_rc(env, rc, "setMask");
  return nullptr;
}
napi_value SetGeometry(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 2) return throw_rc(env, -1, "setGeometry");
  napi_value g = a.v[1];
  napi_value vm = prop(env, g, "mask"), vA = prop(env, g, "geoA"), vG = prop(env, g, "geoG"), vM = prop(env, g, "geoM");
  napi_value vJ = prop(env, g, "jWall");
  if (!vm || !vA || !vG || !vM) return throw_rc(env, -1, "setGeometry: geo needs mask, geoA, geoG, geoM");
  float* jw = nullptr;
  if (vJ) {
    napi_valuetype t;
    napi_typeof(env, vJ, &t);
    if (t == napi_object) jw = typed_data<float>(env, vJ, nullptr);
  }
  int rc = nz_set_geometry(solver_of(env, a.v[0]), typed_data<uint8_t>(env, vm, nullptr), typed_data<float>(env, vA, nullptr),
                           typed_data<float>(env, vG, nullptr), typed_data<float>(env, vM, nullptr), jw);
  if (rc) return throw_rc(env, rc, "setGeometry");
  return nullptr;
}
napi_value SetState(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 3) return throw_rc(env, -1, "setState");
  if (int rc = nz_set_state(solver_of(env, a.v[0]), typed_data<float>(env, a.v[1], nullptr), typed_data<float>(env, a.v[2], nullptr)))
    return throw_rc(env, rc, "setState");
  return nullptr;
}
napi_value Step(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 3) return throw_rc(env, -1, "step");
  nz_params p;
  params_from_js(env, a.v[2], p);
  if (int rc = nz_step(solver_of(env, a.v[0]), (int)num(env, a.v[1]), &p)) return throw_rc(env, rc, "step");
  return nullptr;
}
napi_value Rescale(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 5) return throw_rc(env, -1, "rescaleAmbient");
  nz_params p;
  params_from_js(env, a.v[4], p);
  if (int rc = nz_rescale_ambient(solver_of(env, a.v[0]), (float)num(env, a.v[1]), (float)num(env, a.v[2]), (float)num(env, a.v[3]), &p))
    return throw_rc(env, rc, "rescaleAmbient");
  return nullptr;
}
napi_value WallRow(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 1) return throw_rc(env, -1, "drawWallRow");
  if (int rc = nz_draw_wall_row(solver_of(env, a.v[0]))) return throw_rc(env, rc, "drawWallRow");
  return nullptr;
}
napi_value ReadRect(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 6) return throw_rc(env, -1, "_readRect");
  const int w = (int)num(env, a.v[4]), h = (int)num(env, a.v[5]);
  if (w <= 0 || h <= 0) return throw_rc(env, -1, "_readRect: empty rectangle");
  std::vector<float> out((size_t)w * h * 4);
  if (int rc = nz_read_rect(solver_of(env, a.v[0]), (int)num(env, a.v[1]), (int)num(env, a.v[2]), (int)num(env, a.v[3]), w, h, out.data()))
    return throw_rc(env, rc, "_readRect");
  return new_f32(env, out.data(), out.size());
}
napi_value Sync(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 1) return throw_rc(env, -1, "sync");
  if (int rc = nz_sync(solver_of(env, a.v[0]))) return throw_rc(env, rc, "sync");
  return nullptr;
}
napi_value AsyncBegin(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 3) return throw_rc(env, -1, "requestAsync");
  uint32_t n = 0;
  NAPI_OK(napi_get_array_length(env, a.v[1], &n));
  std::vector<nz_read_req> reqs(n);
  for (uint32_t k = 0; k < n; ++k) {
    napi_value r;
    NAPI_OK(napi_get_element(env, a.v[1], k, &r));
    reqs[k].field = (int)pnum(env, r, "field");
    reqs[k].x = (int)pnum(env, r, "x");
    reqs[k].y = (int)pnum(env, r, "y");
    reqs[k].w = (int)pnum(env, r, "w");
    reqs[k].h = (int)pnum(env, r, "h");
  }
  const int rc = nz_read_async_begin(solver_of(env, a.v[0]), (int)n, reqs.data(), (int32_t)num(env, a.v[2]));
  if (rc < 0) return throw_rc(env, rc, "requestAsync");
  return mk_bool(env, rc == 1);
}
napi_value AsyncPoll(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 2) return throw_rc(env, -1, "pollAsync");
  const size_t floats = (size_t)num(env, a.v[1]);
  std::vector<float> out(floats > 0 ? floats : 1);
  int32_t tag = 0;
  const int rc = nz_read_async_poll(solver_of(env, a.v[0]), out.data(), &tag);
  if (rc < 0) return throw_rc(env, rc, "pollAsync");
  if (rc == 0) return mk_null(env);
  napi_value res;
  napi_create_object(env, &res);
  set(env, res, "tag", mk_num(env, tag));
  set(env, res, "data", new_f32(env, out.data(), floats));
  return res;
}
napi_value AsyncCancel(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 1) return nullptr;
  nz_read_async_cancel(solver_of(env, a.v[0]));
  return nullptr;
}
napi_value SprayInjector(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 2) return throw_rc(env, -1, "setInjector");
  if (int rc = nz_spray_set_injector(solver_of(env, a.v[0]), typed_data<float>(env, a.v[1], nullptr))) return throw_rc(env, rc, "setInjector");
  return nullptr;
}
napi_value SprayReset(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 1) return throw_rc(env, -1, "spray.reset");
  if (int rc = nz_spray_reset(solver_of(env, a.v[0]))) return throw_rc(env, rc, "spray.reset");
  return nullptr;
}
napi_value SprayCount(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 2) return throw_rc(env, -1, "spray.setCount");
  if (int rc = nz_spray_set_count(solver_of(env, a.v[0]), (int)num(env, a.v[1]))) return throw_rc(env, rc, "spray.setCount");
  return nullptr;
}
napi_value SprayStats(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 1) return throw_rc(env, -1, "spray.stats");
  nz_spray_stats_t st;
  if (int rc = nz_spray_stats(solver_of(env, a.v[0]), &st)) return throw_rc(env, rc, "spray.stats");
  napi_value o;
  napi_create_object(env, &o);
  set(env, o, "fuel", mk_num(env, st.fuel));
  set(env, o, "ox", mk_num(env, st.ox));
  set(env, o, "parcels", mk_num(env, (double)st.parcels));
  set(env, o, "meanD", mk_num(env, st.meanD));
  return o;
}
napi_value Probes(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 6) return throw_rc(env, -1, "setProbes");
  nz_probes p{};
  p.iExit = (int32_t)num(env, a.v[1]);
  p.jExit = (int32_t)num(env, a.v[2]);
  p.iThroat = (int32_t)num(env, a.v[3]);
  p.iChamber = (int32_t)num(env, a.v[4]);
  p.dx = num(env, a.v[5]);
  // (pa, Rs0, Rs1, Rs2, RaN, sprayMdot): the reference's P values in double
  if (a.n >= 12) {
    p.exact = 1;
    p.pa = num(env, a.v[6]);
    for (int k = 0; k < 3; ++k) p.Rs[k] = num(env, a.v[7 + k]);
    p.RaN = num(env, a.v[10]);
    p.sprayMdot = num(env, a.v[11]);
  }
  if (int rc = nz_set_probes(solver_of(env, a.v[0]), &p)) return throw_rc(env, rc, "setProbes");
  return nullptr;
}
napi_value Integrals(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 2) return throw_rc(env, -1, "integrals");
  nz_params p;
  params_from_js(env, a.v[1], p);
  nz_integrals r;
  if (int rc = nz_integrals_read(solver_of(env, a.v[0]), &p, &r)) return throw_rc(env, rc, "integrals");
  napi_value o;
  napi_create_object(env, &o);
  set(env, o, "F", mk_num(env, r.F));
  set(env, o, "mdot", mk_num(env, r.mdot));
  set(env, o, "Me", mk_num(env, r.Me));
  set(env, o, "pe", mk_num(env, r.pe));
  set(env, o, "Isp", mk_num(env, r.Isp));
  set(env, o, "mdotThroat", mk_num(env, r.mdotThroat));
  set(env, o, "Pc", mk_num(env, r.Pc));
  set(env, o, "Tc", mk_num(env, r.Tc));
  set(env, o, "mass", mk_num(env, r.mass));
  set(env, o, "energy", mk_num(env, r.energy));
  set(env, o, "residual", mk_num(env, r.residual));
  set(env, o, "unburntOx", mk_num(env, r.unburntOx));
  set(env, o, "fuelVapour", mk_num(env, r.fuelVapour));
  return o;
}
// render.js derive(V) (render.js:420-435): V carries fieldIndex, range, RaN, Rs, chemK,
// Tc, Ta, schlGain, dx as app.js V() builds it (app.js:375-382)
napi_value Derived(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 4) return throw_rc(env, -1, "derive");
  nz_solver* s = solver_of(env, a.v[0]);
  const int Nx = (int)num(env, a.v[2]), Nr = (int)num(env, a.v[3]);
  if (!s || Nx <= 0 || Nr <= 0) return throw_rc(env, -1, "derive");
  nz_derived_params d;
  std::memset(&d, 0, sizeof(d));
  napi_value V = a.v[1];
  d.field = (int32_t)pnum(env, V, "fieldIndex");
  parr(env, V, "range", d.range, 2);
  d.RaN = (float)pnum(env, V, "RaN");
  parr(env, V, "Rs", d.Rs, 3);
  parr(env, V, "chemK", d.chemK, 4);
  d.Tc = (float)pnum(env, V, "Tc");
  d.Ta = (float)pnum(env, V, "Ta");
  d.schlGain = (float)pnum(env, V, "schlGain");
  d.dx = (float)pnum(env, V, "dx");
  std::vector<float> out((size_t)Nx * Nr * 4);
  if (int rc = nz_read_derived(s, &d, out.data())) return throw_rc(env, rc, "derive");
  return new_f32(env, out.data(), out.size());
}
// wallSeparation(setup, row, paSi) (simsetup.js:352-373): object or null
napi_value WallSeparation(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 3) return throw_rc(env, -1, "wallSeparation");
  nz_setup* st = setup_of(env, a.v[0]);
  size_t n = 0;
  const float* row = typed_data<float>(env, a.v[1], &n);
  nz_setup_info in;
  if (!st || !row || nz_setup_get_info(st, &in) || n < (size_t)in.Nx * 4) return throw_rc(env, -1, "wallSeparation");
  nz_separation r;
  if (nz_wall_separation(st, row, num(env, a.v[2]), &r) != 1) return mk_null(env);
  napi_value o;
  napi_create_object(env, &o);
  set(env, o, "x", mk_num(env, r.x));
  set(env, o, "frac", mk_num(env, r.frac));
  set(env, o, "eps", mk_num(env, r.eps));
  set(env, o, "pRatio", mk_num(env, r.pRatio));
  set(env, o, "Msep", mk_num(env, r.Msep));
  return o;
}
// summerfieldPoint(contour, Pc, g, pa) (simsetup.js:375-384): fraction or null
napi_value SummerfieldPoint(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 4) return throw_rc(env, -1, "summerfieldPoint");
  nz_setup* st = setup_of(env, a.v[0]);
  if (!st) return throw_rc(env, -1, "summerfieldPoint");
  double f = 0;
  if (nz_summerfield_point(st, num(env, a.v[1]), num(env, a.v[2]), num(env, a.v[3]), &f) != 1) return mk_null(env);
  return mk_num(env, f);
}
napi_value DeviceAvailable(napi_env env, napi_callback_info) { return mk_bool(env, nz_device_available() == 1); }

// ------------------------------------------------------------------ host setup (buildSimSetup)
void cfg_from_js(napi_env env, napi_value o, nz_config& c) {
  nz_config_default(&c);
  auto d = [&](const char* n, double& f) { napi_value v = prop(env, o, n); if (v) f = num(env, v, f); };
  auto i = [&](const char* n, int32_t& f) { napi_value v = prop(env, o, n); if (v) f = (int32_t)num(env, v, f); };
  auto s = [&](const char* n) -> std::string {
    napi_value v = prop(env, o, n);
    if (!v) return "";
    char buf[64] = {0};
    size_t len = 0;
    if (napi_get_value_string_utf8(env, v, buf, sizeof(buf), &len) != napi_ok) return "";
    return std::string(buf, len);
  };
  d("Dt", c.Dt); d("De", c.De); d("bellPct", c.bellPct); d("coneAngle", c.coneAngle); d("CR", c.CR); d("chamberLen", c.chamberLen);
  d("Pc", c.Pc); d("Tc", c.Tc); d("gamma", c.gamma); d("Mw", c.Mw); d("alt", c.alt); d("mach", c.mach);
  i("base", c.base); d("bodyRatio", c.bodyRatio); d("domainL", c.domainL); i("Nx", c.Nx); d("aspect", c.aspect);
  d("rampMs", c.rampMs); d("cfl", c.cfl); i("order2", c.order2); i("minCells", c.minCells); i("viscous", c.viscous);
  i("afterburn", c.afterburn); d("qAb", c.qAb); d("sAir", c.sAir); d("Tw", c.Tw); d("inletAbsorb", c.inletAbsorb);
  i("injMode", c.injMode); i("injRings", c.injRings); d("injDP", c.injDP); d("injD", c.injD); d("injTMR", c.injTMR);
  d("injAngle", c.injAngle); d("OF", c.OF); d("film", c.film); d("baffles", c.baffles); d("baffleH", c.baffleH);
  d("cavD", c.cavD); d("cavW", c.cavW); d("dx", c.dx); d("hllBlend", c.hllBlend); d("convAngle", c.convAngle);
  d("ambP", c.ambP); d("ambT", c.ambT); i("turbulence", c.turbulence);
  const std::string ty = s("type"), sm = s("startMode"), wm = s("wallModel"), pr = s("prop"), it = s("injType");
  if (!ty.empty()) c.type = ty == "cone" ? NZ_CONE : NZ_BELL;
  if (!sm.empty()) c.startMode = sm == "burst" ? NZ_START_BURST : NZ_START_RAMP;
  if (!wm.empty()) c.wallModel = wm == "stair" ? NZ_WALL_STAIR : NZ_WALL_CUT;
  const char* props[] = {"kerolox", "methalox", "hydrolox", "hybrid", "solid", "hydrazine", "coldgas", "custom"};
  for (int k = 0; k < 8; ++k) if (pr == props[k]) c.prop = k;
  const char* inj[] = {"shower", "coax", "impinge", "pintle"};
  for (int k = 0; k < 4; ++k) if (it == inj[k]) c.injType = k;
}
napi_value BuildSetup(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 1) return throw_rc(env, -1, "buildSimSetup");
  nz_config c;
  cfg_from_js(env, a.v[0], c);
  nz_setup* s = nullptr;
  if (int rc = nz_setup_build(&c, &s)) return throw_rc(env, rc, "buildSimSetup");
  nz_setup_info in;
  nz_setup_get_info(s, &in);
  const size_t N = (size_t)in.Nx * in.Nr, NL = (size_t)(in.Nx + 1) * (in.Nr + 1);
  std::vector<uint8_t> mask(N);
  std::vector<float> geoA(NL * 4), geoG(N * 4), geoM(N * 4), jWall((size_t)in.Nx * 4), kappa(N), Ax(NL), Ar(NL);
  nz_setup_get_geometry(s, mask.data(), geoA.data(), geoG.data(), geoM.data(), jWall.data(), kappa.data(), Ax.data(), Ar.data());
  std::vector<float> A(N * 4), B(N * 4), tab(256 * 12);
  nz_setup_initial_state(s, A.data(), B.data());
  nz_params P;
  nz_setup_get_params(s, &P);
  const bool injector = nz_setup_get_injector(s, tab.data()) == 0;
  napi_value out, geo, init;
  napi_create_object(env, &out);
  napi_create_object(env, &geo);
  napi_create_object(env, &init);
  set(env, geo, "mask", new_u8(env, mask.data(), N));
  set(env, geo, "geoA", new_f32(env, geoA.data(), geoA.size()));
  set(env, geo, "geoG", new_f32(env, geoG.data(), geoG.size()));
  set(env, geo, "geoM", new_f32(env, geoM.data(), geoM.size()));
  set(env, geo, "jWall", new_f32(env, jWall.data(), jWall.size()));
  set(env, geo, "kappa", new_f32(env, kappa.data(), kappa.size()));
  set(env, geo, "Ax", new_f32(env, Ax.data(), Ax.size()));
  set(env, geo, "Ar", new_f32(env, Ar.data(), Ar.size()));
  set(env, geo, "W1", mk_num(env, in.Nx + 1));
  set(env, out, "geo", geo);
  set(env, out, "mask", new_u8(env, mask.data(), N));
  set(env, init, "A", new_f32(env, A.data(), A.size()));
  set(env, init, "B", new_f32(env, B.data(), B.size()));
  set(env, out, "init", init);
  set(env, out, "P", params_to_js(env, P, nullptr));
  // the double-precision P values the integrated quantities use (as the reference's P)
  nz_probes pr{};
  nz_setup_get_probes(s, &pr);
  napi_value pq;
  napi_create_object(env, &pq);
  set(env, pq, "pa", mk_num(env, pr.pa));
  napi_value rs;
  napi_create_array_with_length(env, 3, &rs);
  for (uint32_t k = 0; k < 3; ++k) napi_set_element(env, rs, k, mk_num(env, pr.Rs[k]));
  set(env, pq, "Rs", rs);
  set(env, pq, "RaN", mk_num(env, pr.RaN));
  set(env, pq, "sprayMdot", mk_num(env, pr.sprayMdot));
  set(env, out, "probes", pq);
  if (injector) {
    napi_value inj;
    napi_create_object(env, &inj);
    set(env, inj, "tab", new_f32(env, tab.data(), tab.size()));
    set(env, out, "injector", inj);
  } else {
    set(env, out, "injector", mk_null(env));
  }
  set(env, out, "Nx", mk_num(env, in.Nx));
  set(env, out, "Nr", mk_num(env, in.Nr));
  set(env, out, "dx", mk_num(env, in.dx));
  set(env, out, "tScale", mk_num(env, in.tScale));
  set(env, out, "iExit", mk_num(env, in.iExit));
  set(env, out, "jExit", mk_num(env, in.jExit));
  set(env, out, "iChamber", mk_num(env, in.iChamber));
  set(env, out, "iThroat", mk_num(env, in.iThroat));
  napi_value th;
  napi_create_object(env, &th);
  set(env, th, "mdot", mk_num(env, in.thMdot));
  set(env, th, "F", mk_num(env, in.thF));
  set(env, th, "Isp", mk_num(env, in.thIsp));
  set(env, th, "pe", mk_num(env, in.thPe));
  set(env, th, "Me", mk_num(env, in.thMe));
  set(env, out, "theory", th);
  napi_value atm;
  napi_create_object(env, &atm);
  set(env, atm, "p", mk_num(env, in.atmP));
  set(env, atm, "T", mk_num(env, in.atmT));
  set(env, out, "atm", atm);
  napi_value ct;
  napi_create_object(env, &ct);
  const std::pair<const char*, double> cf[] = {{"Rt", in.Rt}, {"Re", in.Re}, {"Rc", in.Rc}, {"Lc", in.Lc}, {"xt", in.xt},
                                               {"xe", in.xe}, {"eps", in.eps}, {"thn", in.thn}, {"the", in.the}};
  for (const auto& kv : cf) set(env, ct, kv.first, mk_num(env, kv.second));
  set(env, out, "contour", ct);
  // the setup stays alive for the HUD read-outs (wallSeparation, summerfieldPoint) and is
  // freed with the JS object that holds it
  napi_value ext;
  NAPI_OK(napi_create_external(env, s, [](napi_env, void* p, void*) { nz_setup_free(static_cast<nz_setup*>(p)); },
                               nullptr, &ext));
  set(env, out, "handle", ext);
  return out;
}

napi_value Init(napi_env env, napi_value exports) {
  struct Fn { const char* name; napi_callback cb; };
  const Fn fns[] = {
      {"create", Create}, {"destroy", Destroy}, {"resize", Resize}, {"setMask", SetMask}, {"setGeometry", SetGeometry},
      {"setState", SetState}, {"step", Step}, {"rescaleAmbient", Rescale}, {"drawWallRow", WallRow}, {"readRect", ReadRect},
      {"sync", Sync}, {"asyncBegin", AsyncBegin}, {"asyncPoll", AsyncPoll}, {"asyncCancel", AsyncCancel},
      {"sprayInjector", SprayInjector}, {"sprayReset", SprayReset}, {"sprayCount", SprayCount}, {"sprayStats", SprayStats},
      {"setProbes", Probes}, {"integrals", Integrals}, {"deviceAvailable", DeviceAvailable}, {"buildSimSetup", BuildSetup},
      {"derive", Derived}, {"wallSeparation", WallSeparation}, {"summerfieldPoint", SummerfieldPoint},
  };
  for (const Fn& f : fns) {
    napi_value fn;
    if (napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn) != napi_ok) return nullptr;
    if (napi_set_named_property(env, exports, f.name, fn) != napi_ok) return nullptr;
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
