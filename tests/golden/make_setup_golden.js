// make_setup_golden.js — writes input/output vectors of the reference's host-side
// setup functions (buildSimSetup, computeParams, buildCutGeometry, initialState,
// buildInjectorTable, nozzleTheory, atmosphere) for a set of configurations, and of the
// integrated quantities (exitPerformance, chamberPerformance, cellPrimSI) on a seeded
// synthetic state of each.
//
// Runs in the development container only, under node, on the reference's own
// source files read from the directory given as argv[2]. Nothing from the
// reference is written out except the numeric results of these functions.
//
// usage: node make_setup_golden.js <reference source dir> <output dir>
'use strict';
const fs = require('fs');
const path = require('path');
const vm = require('vm');

const srcDir = process.argv[2];
const outDir = process.argv[3];
fs.mkdirSync(outDir, { recursive: true });

// the reference concatenates its modules into one page script (source/build.py);
// the host-side ones are loaded here the same way, into one context
const files = ['physics.js', 'cutcell.js', 'simsetup.js'];
let code = files.map((f) => fs.readFileSync(path.join(srcDir, f), 'utf8')).join('\n;\n');
code += '\n;this.__api = { buildSimSetup, computeParams, buildContour, nozzleTheory, atmosphere, exitPerformance, chamberPerformance, cellPrimSI };\n';
const ctx = vm.createContext({ Math, Float32Array, Float64Array, Uint8Array, Int32Array, Array, Object, JSON, Infinity, NaN, isFinite });
vm.runInContext(code, ctx, { filename: 'reference-host.js' });
const api = ctx.__api;

// cfg as cfgFromS builds it (app.js:201-211) from S defaults (app.js:12-26)
function baseCfg() {
  return {
    Dt: 0.023, De: 0.119, type: 'bell', bellPct: 80, coneAngle: 15, CR: 6, chamberLen: 1.6, Pc: 196e5, Tc: 3670,
    gamma: 1.22, Mw: 22, alt: 0, mach: 0, base: false, bodyRatio: 1.5, domainL: 1.2, Nx: 1024, aspect: 0.28,
    startMode: 'ramp', rampMs: 0.12, cfl: 0.6, order2: true, wallModel: 'cut', minCells: 8, viscous: true,
    afterburn: true, qAb: 6.8, sAir: 1.8, Tw: 800, inletAbsorb: 1, prop: 'kerolox', injMode: false, injType: 'shower',
    injRings: 6, injDP: 20, injD: 1.5, injTMR: 1, injAngle: 30, OF: 2.3, film: 0, baffles: 0, baffleH: 30, cavD: 0, cavW: 40,
  };
}
const PROPS = {
  kerolox: { Tc: 3670, gamma: 1.22, Mw: 22.0, qAb: 6.8, sAir: 1.8, OF: 2.3 },
  methalox: { Tc: 3550, gamma: 1.21, Mw: 20.5, qAb: 3.0, sAir: 0.8, OF: 3.0 },
  hydrolox: { Tc: 3400, gamma: 1.20, Mw: 13.5, qAb: 4.2, sAir: 1.2, OF: 6.0 },
  coldgas: { Tc: 300, gamma: 1.40, Mw: 28.0, qAb: 0, sAir: 1 },
};
function withProp(c, p) { const q = PROPS[p]; return Object.assign(c, { prop: p, Tc: q.Tc, gamma: q.gamma, Mw: q.Mw, qAb: q.qAb, sAir: q.sAir }, q.OF ? { OF: q.OF } : {}); }
const HW = { film: 0, baffles: 0, baffleH: 30, cavD: 0, cavW: 40 };
function inj(c, o) { return Object.assign(c, { injMode: true }, HW, o); }

const cases = {
  // cold-gas cone, 10 mm throat, area ratio 4, 10 bar / 300 K, 0.3 m domain, 256 x 96
  cold_cone_256: Object.assign(withProp(baseCfg(), 'coldgas'), { Dt: 0.020, De: 0.040, type: 'cone', coneAngle: 15, CR: 6.25, chamberLen: 1.4, Pc: 10e5, Tc: 300, domainL: 0.3, Nx: 256, aspect: 0.375 }),
  demo_384: Object.assign(baseCfg(), { Nx: 384 }),
  diamonds_256: Object.assign(withProp(baseCfg(), 'methalox'), { Pc: 90e5, Dt: 0.030, De: 0.125, CR: 5, chamberLen: 1.5, domainL: 1.2, Nx: 256 }),
  hot_bell_192: Object.assign(withProp(baseCfg(), 'methalox'), { Pc: 70e5, Tc: 3500, gamma: 1.2, Mw: 8314.462618 / 380, Dt: 0.040, De: 0.200, CR: 5, chamberLen: 1.5, domainL: 2.2, Nx: 192, aspect: 1 / 3, qAb: 3.0, sAir: 0.8 }),
  flight_base_256: Object.assign(baseCfg(), { Pc: 100e5, Dt: 0.026, De: 0.090, CR: 5, chamberLen: 1.5, alt: 12000, mach: 2.0, base: true, domainL: 1.2, Nx: 256 }),
  stair_192: Object.assign(baseCfg(), { Nx: 192, wallModel: 'stair' }),
  burst_160: Object.assign(withProp(baseCfg(), 'hydrolox'), { Pc: 120e5, Dt: 0.020, De: 0.160, CR: 5, chamberLen: 1.4, domainL: 1.4, Nx: 160, startMode: 'burst' }),
  inj_impinge_192: inj(Object.assign(baseCfg(), { Pc: 70e5, Dt: 0.030, De: 0.085, CR: 5, chamberLen: 6, domainL: 0.5, Nx: 192 }), { injType: 'impinge', injRings: 4, injDP: 20, injD: 1.2, injAngle: 30 }),
  inj_coax_160: inj(Object.assign(withProp(baseCfg(), 'methalox'), { Pc: 70e5, Dt: 0.030, De: 0.085, CR: 5, chamberLen: 6, domainL: 0.5, Nx: 160 }), { injType: 'coax', injRings: 4, injDP: 15, injD: 2 }),
  inj_shower_160: inj(Object.assign(baseCfg(), { Pc: 70e5, Dt: 0.030, De: 0.085, CR: 5, chamberLen: 6, domainL: 0.5, Nx: 160 }), { injType: 'shower', injRings: 6, injDP: 20, injD: 1.5 }),
  inj_pintle_160: inj(Object.assign(baseCfg(), { Pc: 70e5, Dt: 0.030, De: 0.085, CR: 5, chamberLen: 6, domainL: 0.5, Nx: 160 }), { injType: 'pintle', injRings: 1, injDP: 20, injD: 1, injTMR: 0.4 }),
  inj_hardware_192: inj(Object.assign(baseCfg(), { Pc: 70e5, Dt: 0.030, De: 0.085, CR: 5, chamberLen: 6, domainL: 0.5, Nx: 192 }), { injType: 'impinge', injRings: 4, injDP: 20, injD: 1.2, injAngle: 30, film: 10, baffles: 1, baffleH: 30, cavD: 12, cavW: 40 }),
  inj_hydro_shower_128: inj(Object.assign(withProp(baseCfg(), 'hydrolox'), { Pc: 70e5, Dt: 0.030, De: 0.085, CR: 5, chamberLen: 6, domainL: 0.5, Nx: 128 }), { injType: 'shower', injRings: 3, injDP: 20, injD: 1.5, film: 5 }),
};

const PKEYS = ['amb', 'res', 'inlet', 'pStart', 'ramp', 'RcCells', 'inj', 'order2', 'hllBlend', 'ambCons', 'ambB', 'gamR', 'uInlet0', 'vMax',
  'cfl', 'dtMax', 'wFloor', 'visc', 'chem', 'gas', 'Rs', 'chamberNu', 'cap', 'chemK', 'Tw'];
function flatParams(P) {
  const out = [];
  for (const k of PKEYS) { const v = P[k]; if (Array.isArray(v)) out.push(...v); else out.push(typeof v === 'boolean' ? (v ? 1 : 0) : v); }
  if (P.spray) out.push(P.spray.mhat, P.spray.tScale); else out.push(0, 0);
  const d = P.injData;
  if (d) out.push(...d.fuelVec, ...d.oxVec, d.eF, d.eO, d.Ge); else for (let k = 0; k < 11; k++) out.push(0);
  return out;
}
function dump(name, arr) {
  const buf = Buffer.from(arr.buffer, arr.byteOffset, arr.byteLength);
  fs.writeFileSync(path.join(outDir, name), buf);
}
// a seeded flow over the setup's grid (32-bit LCG, Numerical Recipes constants): density
// and pressure jittered about the initial state, a streamwise velocity of up to 1.3
// U_REF, radial velocity, and exhaust / fuel / oxidiser fractions, stored as Float32 the
// way the solver's textures hold them
function syntheticState(s, init, seed) {
  let x = seed >>> 0;
  const rnd = () => { x = (Math.imul(x, 1664525) + 1013904223) >>> 0; return x / 4294967296; };
  const N = s.Nx * s.Nr, A = new Float32Array(N * 4), B = new Float32Array(N * 4);
  for (let k = 0; k < N; k++) {
    const q = k * 4, G = init.B[q + 2] > 0.5 ? init.B[q + 2] : 2.5;
    const r0 = Math.max(init.A[q], 1e-3), p0 = Math.max((init.A[q + 3] - 0.5 * (init.A[q + 1] ** 2 + init.A[q + 2] ** 2) / r0) / G, 1e-3);
    const rho = r0 * (0.8 + 0.4 * rnd()), u = 1.5 * rnd() - 0.2, v = 0.3 * (rnd() - 0.5), p = p0 * (0.8 + 0.4 * rnd());
    const Z = rnd(), F = Z * rnd(), O = (1 - Z) * rnd();
    A[q] = rho; A[q + 1] = rho * u; A[q + 2] = rho * v; A[q + 3] = p * G + 0.5 * rho * (u * u + v * v);
    B[q] = rho * Z; B[q + 1] = rho * F; B[q + 2] = G; B[q + 3] = rho * O;
  }
  return { A, B };
}
function column(s, st, i) {
  const A = new Float32Array(s.Nr * 4), B = new Float32Array(s.Nr * 4);
  for (let j = 0; j < s.Nr; j++) for (let c = 0; c < 4; c++) { A[j * 4 + c] = st.A[(j * s.Nx + i) * 4 + c]; B[j * 4 + c] = st.B[(j * s.Nx + i) * 4 + c]; }
  return { A, B };
}
const index = {};
let caseSeed = 20240501;
for (const [name, cfg] of Object.entries(cases)) {
  const s = api.buildSimSetup(JSON.parse(JSON.stringify(cfg)));
  const init = s.initialState();
  dump(`${name}.mask.u8`, s.geo.mask);
  dump(`${name}.geoA.f32`, s.geo.geoA);
  dump(`${name}.geoG.f32`, s.geo.geoG);
  dump(`${name}.geoM.f32`, s.geo.geoM);
  dump(`${name}.jWall.f32`, s.geo.jWall);
  dump(`${name}.kappa.f32`, s.geo.kappa);
  dump(`${name}.A.f32`, init.A);
  dump(`${name}.B.f32`, init.B);
  if (s.injector) dump(`${name}.tab.f32`, s.injector.tab);
  dump(`${name}.P.f64`, new Float64Array(flatParams(s.P)));
  // live re-parameterisation with other chamber pressure / start pressure (applyLive, app.js:285-304)
  const P2 = api.computeParams(cfg, s.contour, s.dx, 0.8 * cfg.Pc, 0.5 * s.atm.p);
  dump(`${name}.P2.f64`, new Float64Array(flatParams(P2)));
  const xs = []; for (let k = 0; k <= 400; k++) xs.push(s.contour.xe * 1.1 * k / 400 - 0.01);
  dump(`${name}.radius.f64`, new Float64Array(xs.map((x) => s.contour.radiusAt(x))));
  dump(`${name}.radius_x.f64`, new Float64Array(xs));
  // integrated quantities of a seeded flow (app.js:534-539, 597)
  // (only what the three functions read is kept: the exit and throat columns, the chamber cell)
  const syn = syntheticState(s, init, caseSeed++);
  const colE = column(s, syn, s.iExit), colT = column(s, syn, s.iThroat);
  const kc = s.iChamber * 4;
  const cellA = syn.A.slice(kc, kc + 4), cellB = syn.B.slice(kc, kc + 4);
  dump(`${name}.exitA.f32`, colE.A); dump(`${name}.exitB.f32`, colE.B);
  dump(`${name}.throatA.f32`, colT.A); dump(`${name}.throatB.f32`, colT.B);
  dump(`${name}.cellA.f32`, cellA); dump(`${name}.cellB.f32`, cellB);
  const integ = {
    exit: api.exitPerformance(s, colE),
    chamber: api.chamberPerformance(s, colT),
    cell: api.cellPrimSI(s, Array.from(cellA), Array.from(cellB)),
  };
  const th = s.theory;
  index[name] = {
    cfg, Nx: s.Nx, Nr: s.Nr, dx: s.dx, tw: s.tw, bodyR: s.bodyR, xBase: s.xBase, tScale: s.tScale,
    iExit: s.iExit, jExit: s.jExit, iChamber: s.iChamber, iThroat: s.iThroat, nCut: s.geo.nCut, nMerged: s.geo.nMerged,
    contour: { Rt: s.contour.Rt, Re: s.contour.Re, Rc: s.contour.Rc, Lc: s.contour.Lc, xt: s.contour.xt, xe: s.contour.xe, eps: s.contour.eps, thn: s.contour.thn, the: s.contour.the },
    atm: s.atm, theory: { Me: th.Me, pe: th.pe, mdot: th.mdot, F: th.F, Isp: th.Isp, cstar: th.cstar, c0: th.c0, uLim: th.uLim },
    injector: !!s.injector, P2args: [0.8 * cfg.Pc, 0.5 * s.atm.p], integ,
    Pd: { pa: s.P.amb[3], Rs: s.P.Rs, RaN: s.P.RaN, inj: s.P.inj, sprayMdot: s.P.spray ? s.P.spray.mdot : 0 },
  };
}
// atmosphere and nozzle theory at a sweep of inputs
const hs = []; for (let h = -500; h <= 125000; h += 2500) hs.push(h);
index.atmosphere = hs.map((h) => [h, api.atmosphere(h)]);
index.theory = [
  [196e5, 3670, 1.22, 22, 0.023, 0.119, 101325], [10e5, 300, 1.4, 28, 0.020, 0.040, 1e5], [70e5, 3500, 1.2, 8314.462618 / 380, 0.04, 0.2, 0.3e5],
  [90e5, 3550, 1.21, 20.5, 0.030, 0.125, 101325], [30e5, 3150, 1.23, 26, 0.030, 0.064, 101325], [100e5, 3400, 1.2, 13.5, 0.016, 0.1, 1197],
].map((a) => [a, api.nozzleTheory({ Pc: a[0], Tc: a[1], gamma: a[2], Mw: a[3], Dt: a[4], De: a[5], pa: a[6] })]);
fs.writeFileSync(path.join(outDir, 'index.json'), JSON.stringify(index));
console.log('wrote', Object.keys(cases).length, 'setup cases to', outDir);
