/*
 * nozzle.h — C ABI of the MI355X-native rocket-nozzle solver hot path.
 *
 * The reference has no C API: its solver is the JavaScript class GPUSolver
 * (source/solver.js:677-1076) plus GPUSpray (source/spray.js:211-321), and its
 * host setup lives in source/simsetup.js, source/cutcell.js, source/physics.js.
 * Every entry point below names the reference method whose behaviour it
 * replaces. The Node-API addon (js/nozzle_napi.cpp) binds these one-to-one
 * under the same method names, so app.js / render.js call sites are unchanged.
 *
 * Conventions
 *  - Plain pointers and sizes. Arrays handed in are copied during the call;
 *    the caller keeps ownership (as texSubImage2D does, solver.js:797-822).
 *  - Arrays handed out are caller-allocated and filled synchronously
 *    (as readPixels does, solver.js:1016-1025).
 *  - Every function returns 0 on success and a negative code on failure;
 *    nz_last_error() then describes the failure (the addon throws a JS Error,
 *    matching the reference's `throw new Error`, solver.js:680,713,723,748).
 *  - Texture-shaped arrays are row-major, k = j*W + i, RGBA-interleaved
 *    float32 exactly as the reference uploads and reads them.
 */
#ifndef NOZZLE_H
#define NOZZLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Parameter block: the reference's normalised `P` object                    */
/* (source/simsetup.js:257-311), as consumed by the GPU passes                */
/* (solver.js:900-907, 914-915, 930-936, 867, 948-981; spray.js:265-271,291). */
/* Floats are the values the reference passes as float32 uniforms.           */
/* ------------------------------------------------------------------------ */
typedef struct nz_params {
  /* boundary / primitive pass (FS_PRIM, solver.js:56-126) */
  float amb[4];        /* free stream primitive (rho,u,v,p) */
  float res[3];        /* reservoir (P0, T0, R_exhaust) normalised */
  float inlet[4];      /* (1/tau_ubar, absorption 0-1, G exhaust, G air) */
  float pStart;        /* pressure at which the ignition ramp starts */
  float ramp;          /* ramp duration (grid time); <= 0: instant */
  float RcCells;       /* chamber radius in cells */
  int32_t inj;         /* 1: injector-resolved chamber */
  /* flux pass (solver.js:914-915) */
  int32_t order2;      /* 1: MUSCL second order */
  float hllBlend;      /* HLL blending sensitivity */
  /* update pass (solver.js:930-936) */
  float ambCons[4];
  float ambB[4];
  float gamR[2];
  float uInlet0;
  float vMax;
  /* CFL (solver.js:867) */
  float cfl, dtMax, wFloor;
  /* physics split step (solver.js:948-981) */
  int32_t visc, chem;
  float gas[4];        /* (R_air, R_exhaust, T_REF, muK) */
  float Rs[3];
  float chamberNu[2];
  float cap[2];
  float chemK[4];
  float Tw;
  /* spray (spray.js:265-271, 291) */
  float sprayMhat, sprayTScale;
  float fuelVec[4], oxVec[4];
  float eF, eO, Ge;
} nz_params;

/* ------------------------------------------------------------------------ */
/* Solver: GPUSolver + GPUSpray                                              */
/* ------------------------------------------------------------------------ */
typedef struct nz_solver nz_solver;

/* readable fields for nz_read_rect (the textures _readRect is used on) */
enum {
  NZ_FIELD_A = 0,        /* state.A  = (rho, rho u, rho v, E)       Nx x Nr */
  NZ_FIELD_B = 1,        /* state.B  = (rho Z, rho F, G, ubar|rhoO) Nx x Nr */
  NZ_FIELD_STEP = 2,     /* stepInfoTex = (dt, t_old, t_new, wmax)   1 x 1  */
  NZ_FIELD_WALLROW = 3,  /* drawWallRow() output (u, v, p, c)        Nx x 1 */
  NZ_FIELD_T2 = 4        /* vis.T2 (mu_eff, h, |S|*sensor, mu_lam)   Nx x Nr */
};

/* last error text of the calling thread ("" if none) */
const char* nz_last_error(void);
/* 1 if a gfx950 HIP device is usable, else 0 (never throws) */
int nz_device_available(void);

/* new GPUSolver(gl) (solver.js:678-704); the spray half is created with it */
int nz_create(nz_solver** out);
/* as nz_create, with the arithmetic type of the state and kernels: 32 (the
 * reference's FP32) or 64 (FP64 run of the same kernels, the error reference) */
int nz_create_ex(nz_solver** out, int precision);
int nz_destroy(nz_solver* s);
/* GPUSolver.resize(Nx, Nr) (solver.js:759-789): frees and reallocates all
 * device buffers, then resets the spray (spray.js:224-235) */
int nz_resize(nz_solver* s, int Nx, int Nr);
int nz_get_size(const nz_solver* s, int* Nx, int* Nr);
/* GPUSolver.setMask(u8) (solver.js:791-799): mask is Nx*Nr, nonzero = solid */
int nz_set_mask(nz_solver* s, const uint8_t* mask);
/* GPUSolver.setGeometry(geo) (solver.js:801-814).
 * geoA: (Nx+1)*(Nr+1)*4, geoG: Nx*Nr*4, geoM: Nx*Nr*4, jWall: Nx*4 or NULL */
int nz_set_geometry(nz_solver* s, const uint8_t* mask, const float* geoA,
                    const float* geoG, const float* geoM, const float* jWall);
/* GPUSolver.setState(A, B) (solver.js:816-828): A, B are Nx*Nr*4; zeroes the
 * step info and simulated time and resets the spray */
int nz_set_state(nz_solver* s, const float* A, const float* B);
/* GPUSolver.step(n, P) (solver.js:849-891): enqueues n steps, returns
 * without waiting for the device */
int nz_step(nz_solver* s, int n, const nz_params* P);
/* Instrumented nz_step (no counterpart in the reference; for benchmarks): the
 * stage-1, spray and stage-2 kernels are launched with start / stop events that
 * record the kernels' own execution (hipExtLaunchKernel); after waiting for the
 * device, ms[0..3] = summed kernel times of stage 1, the spray, stage 2, and the
 * span of the n steps on the solver's stream. */
int nz_step_timed(nz_solver* s, int n, const nz_params* P, double* ms);
/* GPUSolver.rescaleAmbient(fp, fT, du, P) (solver.js:998-1010) */
int nz_rescale_ambient(nz_solver* s, float fp, float fT, float du, const nz_params* P);
/* Tile shapes of the fused stage kernels (no counterpart in the reference, whose
 * passes are full-screen draws). nz_tile_shapes fills tx[k], ty[k] and waves[k] (the
 * waves per SIMD the kernels without the split step are compiled to fit; 1 = no
 * constraint) for k < cap, any pointer may be NULL, and returns how many shapes the
 * library holds at this precision; shape 0 is the default. The shape changes speed only, never results beyond FP rounding order
 * (none: each face and cell is evaluated by the same arithmetic in every shape). */
int nz_tile_shapes(int precision, int* tx, int* ty, int* waves, int cap);
int nz_set_tile_shape(nz_solver* s, int shape);
int nz_get_tile_shape(const nz_solver* s, int* shape);
/* Autotuner: times `steps` steps of every shape from the current state, restoring
 * the state (fields, parcels, step info) after each trial, and keeps the fastest.
 * us_per_step (NULL or >= nz_tile_shapes() entries) receives each shape's time. */
int nz_autotune(nz_solver* s, const nz_params* P, int steps, int* best, double* us_per_step);
/* GPUSolver.drawWallRow() (solver.js:986-996): fills NZ_FIELD_WALLROW */
int nz_draw_wall_row(nz_solver* s);
/* GPUSolver._readRect(tex, x, y, w, h) (solver.js:1016-1025):
 * out holds w*h*4 floats */
int nz_read_rect(nz_solver* s, int field, int x, int y, int w, int h, float* out);
/* GPUSolver.readStepInfo() (solver.js:1072): out[4] */
int nz_read_step_info(nz_solver* s, float* out);
/* the whole state at the solver's precision, as doubles (no counterpart in the
 * reference; the FP64 error reference reads through it): A, B are Nx*Nr*4 */
int nz_read_state_f64(nz_solver* s, double* A, double* B);
/* number of completed steps since setState (GPUSolver.stepCount) */
int nz_step_count(const nz_solver* s, int64_t* out);
/* block until all enqueued work is finished */
int nz_sync(nz_solver* s);

/* asynchronous read-back (requestAsync / pollAsync / cancelAsync,
 * solver.js:1028-1070). Requests are packed back to back into one buffer. */
typedef struct nz_read_req { int32_t field, x, y, w, h; } nz_read_req;
/* returns 1 if started, 0 if a read is still in flight (requestAsync false) */
int nz_read_async_begin(nz_solver* s, int nreq, const nz_read_req* reqs, int32_t tag);
/* returns 1 and fills out (sum of w*h*4 floats) + *tag when done, 0 if not */
int nz_read_async_poll(nz_solver* s, float* out, int32_t* tag);
int nz_read_async_cancel(nz_solver* s);
/* 1 while a request is outstanding (the reference's `_async` field) */
int nz_read_async_pending(const nz_solver* s);

/* GPUSpray */
typedef struct nz_spray_stats_t { double fuel, ox; int64_t parcels; double meanD; } nz_spray_stats_t;
/* GPUSpray.setInjector({tab}) (spray.js:236-241): tab is 256*3*4 floats */
int nz_spray_set_injector(nz_solver* s, const float* tab);
/* GPUSpray.reset() (spray.js:242-254) */
int nz_spray_reset(nz_solver* s);
/* GPUSpray.stats(P) (spray.js:312-320) */
int nz_spray_stats(nz_solver* s, nz_spray_stats_t* out);
/* Parcel count. The reference fixes 256*256 = 65,536 (spray.js:16); a
 * multiple of 4 between 4 and 2^22 is accepted. Resets the spray. */
int nz_spray_set_count(nz_solver* s, int nparcels);
/* parcel read-back for tests: PA, PB each nparcels*4 floats (spray.js:4-7) */
int nz_spray_read(nz_solver* s, float* PA, float* PB);

/* Derived display fields, FS_DERIVED (render.js:60-134) on the device: per cell
 * (t, exhaust fraction Z, normalised temperature, Mach) of the current state, with
 * solid cells (0, 0, 0, -1). The view parameters are those render.js derive() passes
 * (render.js:420-435, from app.js:379-382). field: 0 T, 1 Mach, 2 schlieren,
 * 3 log10 p, 4 log10 rho, 5 speed, 6 vorticity, 7 Z, 8 (as 0), 9 log10 mu_eff/mu,
 * 10 burnt fraction, 11 F, 12 O, 13 burn rate. Fields 9 and 13 use vis.T2, which
 * the library recomputes from the last stage-1 state (as NZ_FIELD_T2). */
typedef struct nz_derived_params {
  int32_t field;
  float range[2];          /* display range (t = (v - range[0]) / (range[1] - range[0])) */
  float RaN;               /* normalised air gas constant (P.gas[0]) */
  float Rs[3];             /* P.Rs */
  float chemK[4];          /* P.chemK */
  float Tc, Ta;            /* chamber and ambient temperature, K */
  float schlGain;          /* schlieren sensitivity */
  float dx;                /* cell size, m */
} nz_derived_params;
int nz_read_derived(nz_solver* s, const nz_derived_params* d, float* out);

/* Integrated quantities, summed on the device in FP64.
 * exitPerformance (simsetup.js:314-337), chamberPerformance (239-250),
 * cellPrimSI at the chamber probe (338-348, app.js:537). */
typedef struct nz_probes {
  int32_t iExit, jExit, iThroat, iChamber;  /* setup.iExit, jExit, iThroat, iChamber */
  double dx;                                /* cell size, m */
  /* exact != 0: the parameter-block values below, at the double precision the reference's
   * P object holds them (simsetup.js:257-311); exact == 0: the float uniforms of the
   * nz_params passed to nz_integrals_read, and a chamberPerformance target of 1 */
  int32_t exact;
  double pa;                                /* P.amb[3], normalised ambient pressure */
  double Rs[3], RaN;                        /* P.Rs, P.RaN (chamber-probe gas constant) */
  double sprayMdot;                         /* P.spray.mdot, kg/s (0 without spray: target 1) */
} nz_probes;
typedef struct nz_integrals {
  double F;            /* exit-plane thrust, N */
  double mdot;         /* exit-plane exhaust mass flow, kg/s (simsetup.js:324) */
  double Me, pe, Isp;  /* exit Mach, mean exit pressure (Pa), Isp (s) */
  double mdotThroat;   /* throat-plane total mass flow, kg/s (chamberPerformance.mdot) */
  double Pc, Tc;       /* chamber probe pressure (Pa), temperature (K) */
  double mass, energy; /* sum over fluid groups of rho*V, E*V (normalised) */
  double residual;     /* L2 of d(rho)/dt over fluid groups, last step, per unit grid time */
  double unburntOx;    /* chamberPerformance.unburntOx: throat-plane rho*O flow / target */
  double fuelVapour;   /* chamberPerformance.fuelVapour: throat-plane rho*F flow / target */
} nz_integrals;
int nz_set_probes(nz_solver* s, const nz_probes* pr);
/* exitPerformance + chamberPerformance + cellPrimSI(iChamber, 0) of the current state
 * (simsetup.js:314-337, 239-250, 338-348; app.js:534-539, 597), summed on the device in
 * FP64 in a fixed order (the same bits on every call) */
int nz_integrals_read(nz_solver* s, const nz_params* P, nz_integrals* out);

/* ------------------------------------------------------------------------ */
/* Batched solves with continuous admission. No counterpart in the          */
/* reference, which runs one GPUSolver per page (app.js:151-167); for       */
/* callers that advance many independent solves (parameter sweeps, a        */
/* request loop). Every resident solve of any grid size advances in the     */
/* same launches (one per kernel per step over all residents' tiles), from  */
/* the same kernels as nz_step, so a solve's result does not depend on what */
/* else is resident (bit-identical, the spray included: its deposit sums   */
/* fixed-point integers, whose order leaves no trace). Jobs are admitted    */
/* first-fit in submission order while the device-memory budget and the     */
/* resident limit allow; their buffers come from one device slab of the     */
/* budget allocated at nz_batch_create. A solve retires when it meets its   */
/* stop test; its memory returns to the slab and queued jobs take its place.*/
/* ------------------------------------------------------------------------ */
typedef struct nz_batch nz_batch;
typedef struct nz_batch_job {
  int32_t Nx, Nr;
  const uint8_t* mask;                     /* as nz_set_geometry */
  const float *geoA, *geoG, *geoM, *jWall; /* jWall may be NULL */
  const float *A, *B;                      /* initial state, as nz_set_state */
  nz_params P;                             /* parameters, held for the whole solve */
  const float* injector;                   /* NULL, or 256*3*4 floats (nz_spray_set_injector) */
  int32_t parcels;                         /* parcel count when an injector is given */
  nz_probes probes;                        /* dx > 0: integrated quantities at retirement */
  double t_end;                            /* stop at this simulated grid time (<= 0: off) */
  double res_tol;                          /* stop at residual <= res_tol (<= 0: off) ... */
  int32_t min_steps;                       /* ... once this many steps are done */
  int32_t max_steps;                       /* stop after this many steps (<= 0: off) */
  int32_t keep_state;                      /* keep the final state for nz_batch_read_state */
} nz_batch_job;
enum { NZ_JOB_QUEUED = 0, NZ_JOB_RUNNING = 1, NZ_JOB_DONE = 2 };
enum { NZ_STOP_NONE = 0, NZ_STOP_RESIDUAL = 1, NZ_STOP_TIME = 2, NZ_STOP_STEPS = 3 };
typedef struct nz_batch_result {
  int32_t state, stop;                     /* NZ_JOB_*, NZ_STOP_* */
  int64_t steps;
  double t_grid, dt, residual;              /* at the last completed step */
  double t_submit, t_admit, t_done;        /* host steady clock, s */
  int32_t have_integrals;
  nz_integrals integ;                      /* at retirement, when probes were given */
} nz_batch_result;
/* precision 32 or 64; hbm_budget in bytes for all residents' buffers;
 * max_resident >= 1 */
int nz_batch_create(nz_batch** out, int precision, uint64_t hbm_budget, int max_resident);
int nz_batch_destroy(nz_batch* b);
/* device bytes one solve of this size holds while resident */
uint64_t nz_batch_instance_bytes(int Nx, int Nr, int precision, int parcels);
/* copy a job into the queue; *id identifies it from then on. Fails if the job
 * alone exceeds the budget. */
int nz_batch_submit(nz_batch* b, const nz_batch_job* job, int64_t* id);
/* admit what fits and enqueue `steps` steps of every resident (a captured graph
 * when enabled); then wait for the chunk the previous call enqueued, read its
 * stop flags and start retiring the solves that stopped in it. The device works
 * on this call's chunk while the host retires and admits, so a run loop keeps it
 * busy. A retiring solve's integrals and kept state are collected at a later call
 * (at once when nothing is left to run), when it is listed by nz_batch_finished.
 * steps = 0: admit only (no waiting, no filing: nz_batch_finished still lists
 * the last stepping call's jobs). *resident counts
 * the solves not yet filed (running or retiring), *queued the waiting jobs (both
 * may be NULL); a loop of run calls ends when both are 0. */
int nz_batch_run(nz_batch* b, int steps, int* resident, int* queued);
/* device seconds of the chunks whose stop flags have been read (event-timed),
 * for the busy fraction of a run loop: busy / wall time */
int nz_batch_gpu_busy(const nz_batch* b, double* seconds);
int nz_batch_result_get(nz_batch* b, int64_t id, nz_batch_result* out);
/* state of a running solve, or of a finished one submitted with keep_state */
int nz_batch_read_state(nz_batch* b, int64_t id, float* A, float* B);
/* forget a finished job (its kept state and result) */
int nz_batch_release(nz_batch* b, int64_t id);
/* 1: capture each run's launches in a HIP graph and replay it while the
 * resident set and step count are unchanged (default 1) */
int nz_batch_set_graph(nz_batch* b, int on);
int nz_batch_memory(const nz_batch* b, uint64_t* used, uint64_t* budget);
/* tile shape of every resident and later admission (nz_tile_shapes) */
int nz_batch_set_tile_shape(nz_batch* b, int shape);
/* autotuner on the live batch: admits what fits, runs `steps` steps (plus one
 * untimed) with every tile shape in turn and keeps the fastest. The solves advance
 * during the trials (it is real work); stop tests still apply and the stopped solves
 * retire at the next nz_batch_run. us_per_step: NULL or >= nz_tile_shapes() entries. */
int nz_batch_autotune(nz_batch* b, int steps, int* best, double* us_per_step);
/* ids of the jobs filed as finished by the last nz_batch_run with steps > 0; returns the count
 * (at most cap are written) */
int nz_batch_finished(const nz_batch* b, int64_t* ids, int cap);

/* ------------------------------------------------------------------------ */
/* Host setup: buildSimSetup / computeParams / buildCutGeometry /            */
/* buildInjectorTable restated in C++ (simsetup.js:7-311, cutcell.js:14-186, */
/* physics.js:7-199). Used to build solver inputs without the browser.       */
/* ------------------------------------------------------------------------ */
enum { NZ_BELL = 0, NZ_CONE = 1 };
enum { NZ_START_RAMP = 0, NZ_START_BURST = 1 };
enum { NZ_WALL_CUT = 0, NZ_WALL_STAIR = 1 };
enum { NZ_PROP_KEROLOX = 0, NZ_PROP_METHALOX, NZ_PROP_HYDROLOX, NZ_PROP_HYBRID,
       NZ_PROP_SOLID, NZ_PROP_HYDRAZINE, NZ_PROP_COLDGAS, NZ_PROP_CUSTOM };
enum { NZ_INJ_SHOWER = 0, NZ_INJ_COAX, NZ_INJ_IMPINGE, NZ_INJ_PINTLE };

/* the `cfg` object of cfgFromS (app.js:201-211), SI units. Fields that the
 * reference reads as `cfg.x || default` treat 0 as "use the default". */
typedef struct nz_config {
  double Dt, De;              /* throat / exit diameter, m */
  int32_t type;               /* NZ_BELL | NZ_CONE */
  double bellPct, coneAngle;  /* %, deg */
  double CR, chamberLen;      /* contraction ratio, chamber length in Rc */
  double Pc, Tc, gamma, Mw;   /* Pa, K, -, kg/kmol */
  double alt, mach;           /* m, - */
  int32_t base; double bodyRatio;
  double domainL;             /* m */
  int32_t Nx; double aspect;
  int32_t startMode; double rampMs;
  double cfl; int32_t order2; int32_t wallModel; int32_t minCells;
  int32_t viscous, afterburn; double qAb, sAir, Tw, inletAbsorb;
  int32_t prop;
  int32_t injMode, injType, injRings;
  double injDP, injD, injTMR, injAngle, OF;
  double film, baffles, baffleH, cavD, cavW;
  double dx;                  /* > 0: keep this cell size (remesh, app.js:259) */
  double hllBlend;            /* cfgFromS leaves it undefined -> 1.0 */
  /* extensions with no counterpart in the reference (defaults keep parity) */
  double convAngle;           /* converging half-angle, deg (reference: 35, physics.js:129) */
  double ambP, ambT;          /* > 0: ambient pressure (Pa) / temperature (K) instead of ISA(alt) */
  int32_t turbulence;         /* 0: eddy-viscosity length scale l = 0 everywhere */
  int32_t deviceGeometry;     /* 1: the cut-cell geometry is built on the GPU (nz_cutcell_build) */
  /* drop-size law of the injector table (0, 0: the reference's per-element sizes, jitter
   * uniform in log over +-0.7, spray.js:53): median drop dropD0Frac x the orifice diameter,
   * lognormal spread dropSigma (written to the table's third row, y) */
  double dropD0Frac, dropSigma;
} nz_config;

typedef struct nz_setup nz_setup;
typedef struct nz_setup_info {
  int32_t Nx, Nr; double dx, tScale, tw, bodyR, xBase;
  int32_t iExit, jExit, iChamber, iThroat;
  int32_t injector;           /* 1 if the injector table is present */
  int32_t nCut, nMerged;
  double Rt, Re, Rc, Lc, xt, xe, eps, thn, the;   /* contour (m, rad) */
  double atmT, atmP;          /* ambient used (K, Pa) */
  /* nozzleTheory (physics.js:75-106) at the ambient pressure */
  double thMe, thPe, thMdot, thF, thIsp, thCstar, thC0, thUlim;
} nz_setup_info;

/* the reference defaults of S -> cfgFromS (app.js:12-26, 201-211) */
void nz_config_default(nz_config* cfg);
/* fill cfg from ENGINE_PRESETS[name] as applyPreset does (app.js:1108-1121) */
int nz_config_preset(nz_config* cfg, const char* name);
int nz_setup_build(const nz_config* cfg, nz_setup** out);
void nz_setup_free(nz_setup* s);
int nz_setup_get_info(const nz_setup* s, nz_setup_info* out);
/* setup.P as float uniforms */
int nz_setup_get_params(const nz_setup* s, nz_params* P);
/* setup.P at full (double) precision, in the field order of nz_params */
int nz_setup_get_params_f64(const nz_setup* s, double* out, int n);
/* computeParams(cfg, contour, dx, PcSi, pStartSi) (simsetup.js:257) for live changes */
int nz_setup_compute_params(const nz_setup* s, const nz_config* cfg, double PcSi,
                            double pStartSi, nz_params* P);
/* geometry arrays as setGeometry takes them; any pointer may be NULL.
 * kappa: Nx*Nr, Ax/Ar: (Nx+1)*(Nr+1) face areas (cutcell.js:185) */
int nz_setup_get_geometry(const nz_setup* s, uint8_t* mask, float* geoA, float* geoG,
                          float* geoM, float* jWall, float* kappa, float* Ax, float* Ar);
/* setup.initialState() (simsetup.js:72-102) */
int nz_setup_initial_state(const nz_setup* s, float* A, float* B);
/* setup.injector.tab (simsetup.js:181-237); returns -1 if not injector mode */
int nz_setup_get_injector(const nz_setup* s, float* tab);
/* probes and double-precision parameter values of the setup for nz_set_probes (exact = 1) */
int nz_setup_get_probes(const nz_setup* s, nz_probes* out);

/* Cut-cell geometry on the GPU: buildCutGeometry (cutcell.js:14-186) as kernels, bit for
 * bit the reference's (and the host restatement's) output. The contour enters as the
 * reference's fine radius table plus its polyline (for the look-up past the table's end). */
typedef struct nz_cutcell_in {
  int32_t Nx, Nr;
  double dx, tw, xBase, bodyR;      /* cell size, wall thickness, base x, body radius (m) */
  int32_t base, stair;              /* base body present; staircase walls (no cut cells) */
  const float* Rtab;                /* Nx*8+1: inner-wall radius (cells) every 1/8 cell, -1e9 past
                                       the exit (cutcell.js:19-27) */
  int32_t npts;
  const double *xs, *rs;            /* contour polyline (m), radiusAt (physics.js:190-197) */
  double Rc, Re, xe, cavD, cavW;    /* chamber / exit radius, exit x, cavity pocket (m) */
  int32_t nsolids;
  const double* solids;             /* nsolids x (x0, x1, r0, r1), m (chamber hardware) */
} nz_cutcell_in;
/* the setup's cut-cell input; its pointers stay valid while the setup lives */
int nz_setup_get_cutcell_input(const nz_setup* s, nz_cutcell_in* out);
/* host arrays in and out, as nz_setup_get_geometry fills them; counts: NULL or (nCut, nMerged).
 * Allocates its device buffers for the call and frees them. */
int nz_cutcell_build(const nz_cutcell_in* in, uint8_t* mask, float* geoA, float* geoG, float* geoM, float* kappa,
                     float* Ax, float* Ar, int32_t* counts);
/* nz_cutcell_build without any device allocation. No counterpart in the reference: for
 * callers that cannot allocate (code under graph capture, serving loops admitting new
 * geometries). Inputs, intermediates and outputs use one 36 MiB device array per GPU, held
 * until the process ends: grids beyond about 1024 x 384 cells (or more than 256 solids) are
 * refused with an error. Runs on the null stream and waits for it. Calls that overlap on
 * one GPU race on that array. Same results as nz_cutcell_build, bit for bit. */
int nz_cutcell_build_noalloc(const nz_cutcell_in* in, uint8_t* mask, float* geoA, float* geoG, float* geoM,
                             float* kappa, float* Ax, float* Ar, int32_t* counts);
/* HUD read-outs on the device's near-wall row (nz_draw_wall_row + NZ_FIELD_WALLROW:
 * Nx x (u, v, p, c), normalised). wallSeparation (simsetup.js:352-373): returns 1
 * and fills *out when the bell has separated, 0 while attached. */
typedef struct nz_separation { double x, frac, eps, pRatio, Msep; } nz_separation;
int nz_wall_separation(const nz_setup* s, const float* row, double paSi, nz_separation* out);
/* summerfieldPoint (simsetup.js:375-384): fraction of the bell where 1-D theory puts
 * the wall pressure below 0.4 pa; returns 1 and fills *frac, or 0 if never */
int nz_summerfield_point(const nz_setup* s, double PcSi, double gamma, double paSi, double* frac);
/* contour radius at x (m) (buildContour.radiusAt, physics.js:190-197) */
double nz_setup_radius_at(const nz_setup* s, double x);
/* nozzleTheory (physics.js:75-106); out = {At,Ae,eps,Me,pe,Te,ue,cstar,mdot,F,Fvac,Isp,c0,uLim} */
int nz_nozzle_theory(double Pc, double Tc, double gamma, double Mw, double Dt, double De,
                     double pa, double* out14);
/* atmosphere(h) (physics.js:35-50); out = {T, p, rho, a} */
void nz_atmosphere(double h, double* out4);

#ifdef __cplusplus
}
#endif
#endif /* NOZZLE_H */
