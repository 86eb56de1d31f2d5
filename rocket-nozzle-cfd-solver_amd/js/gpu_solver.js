// gpu_solver.js — drop-in GPUSolver / GPUSpray for the reference's JavaScript host
// (source/solver.js:677-1076, source/spray.js:211-321), running the MI355X HIP
// kernels through the Node-API addon build_out/nozzle_napi.node.
//
// Same constructor, methods, arguments and return shapes as the reference, so the
// call sites in app.js (restart 229-252, remesh 254-270, applyLive 285-304,
// frame 953-992, readback 534-539, asyncReadback 576-603, self-test 1246-1427) and
// render.js (derive/display 420-503) are unchanged. Textures are named handles:
// state.A / state.B, stepInfoTex, drawWallRow()'s row, vis.T2 — _readRect and
// requestAsync accept them wherever the reference passes a WebGL texture.
'use strict';
const path = require('path');

const addon = require(path.join(__dirname, '..', 'build_out', 'nozzle_napi.node'));

const FIELD = { A: 0, B: 1, step: 2, wallRow: 3, T2: 4 };
const tex = (name) => ({ __nzField: FIELD[name], name });

class GPUSolver {
  // `gl` is accepted for call-site compatibility and ignored; `opts.precision` 32 | 64
  constructor(gl, opts = {}) {
    this.gl = gl || null;
    this._h = addon.create(opts.precision || 32);
    this.Nx = 0; this.Nr = 0;
    this.stepCount = 0;
    this._async = null;
    this.mask = null;
    this.vis = { T2: tex('T2') };
    this.prim = { W: null, Wb: null };   // the padded primitives live in LDS only
  }
  resize(Nx, Nr) {
    this.cancelAsync();
    addon.resize(this._h, Nx, Nr);
    this.Nx = Nx; this.Nr = Nr;
    if (this.spray) this.spray.resize();
  }
  setMask(u8) { addon.setMask(this._h, u8); this.maskData = u8; }
  setGeometry(geo) {
    addon.setGeometry(this._h, geo);
    this.maskData = geo.mask;
    this.geoM = geo.geoM; this.geoG = geo.geoG;
  }
  setState(A, B) {
    addon.setState(this._h, A, B);
    this.stepCount = 0;
    if (this.spray) this.spray.reset();
  }
  step(n, P) {
    addon.step(this._h, n, P);
    this.stepCount += n;
  }
  drawWallRow() { addon.drawWallRow(this._h); return tex('wallRow'); }
  rescaleAmbient(fp, fT, du, P) { addon.rescaleAmbient(this._h, fp, fT, du, P); }
  get state() { return { A: tex('A'), B: tex('B') }; }
  get stepInfoTex() { return tex('step'); }
  _readRect(t, x, y, w, h) { return addon.readRect(this._h, t.__nzField, x, y, w, h); }
  requestAsync(reqs, tag) {
    if (this._async) return false;
    let floats = 0;
    const offs = reqs.map((r) => { const o = floats; floats += r.w * r.h * 4; return o; });
    const ok = addon.asyncBegin(this._h, reqs.map((r) => ({ field: r.tex.__nzField, x: r.x, y: r.y, w: r.w, h: r.h })), tag);
    if (!ok) return false;
    this._async = { reqs, offs, floats, tag };
    return true;
  }
  pollAsync() {
    const a = this._async;
    if (!a) return null;
    const res = addon.asyncPoll(this._h, a.floats);
    if (!res) return null;
    this._async = null;
    const data = {};
    a.reqs.forEach((r, k) => { data[r.name] = res.data.subarray(a.offs[k], a.offs[k] + r.w * r.h * 4); });
    return { tag: a.tag, data, probe: a.probe };
  }
  cancelAsync() {
    if (!this._async) return;
    addon.asyncCancel(this._h);
    this._async = null;
  }
  readStepInfo() { return this._readRect(this.stepInfoTex, 0, 0, 1, 1); }
  readColumn(i) { return { A: this._readRect(this.state.A, i, 0, 1, this.Nr), B: this._readRect(this.state.B, i, 0, 1, this.Nr) }; }
  readCell(i, j) { return { A: this._readRect(this.state.A, i, j, 1, 1), B: this._readRect(this.state.B, i, j, 1, 1) }; }
  readAll() { return { A: this._readRect(this.state.A, 0, 0, this.Nx, this.Nr), B: this._readRect(this.state.B, 0, 0, this.Nx, this.Nr) }; }
  // build-defined extras (no counterpart in the reference)
  sync() { addon.sync(this._h); }
  // the probes of exitPerformance / chamberPerformance / cellPrimSI, with the setup's P values
  // in double (simsetup.js:239-348) when it carries them
  setProbes(setup) {
    const P = setup.P;
    const q = setup.probes || (P && P.amb && P.Rs
      ? { pa: P.amb[3], Rs: P.Rs, RaN: P.RaN === undefined ? 1 : P.RaN, sprayMdot: P.spray ? P.spray.mdot : 0 } : null);
    if (q) {
      addon.setProbes(this._h, setup.iExit, setup.jExit, setup.iThroat, setup.iChamber, setup.dx, q.pa,
        q.Rs[0], q.Rs[1], q.Rs[2], q.RaN, q.sprayMdot);
    } else {
      addon.setProbes(this._h, setup.iExit, setup.jExit, setup.iThroat, setup.iChamber, setup.dx);
    }
  }
  integrals(P) { return addon.integrals(this._h, P); }
  // render.js derive(V, second) (render.js:420-435): the derived field's Nx x Nr x
  // (t, Z, Tn, M) values, returned instead of drawn into derivedFbo
  derive(V, second = false) {
    if (second) V = Object.assign({}, V, { fieldIndex: V.fieldIndex2, range: V.range2 });
    return addon.derive(this._h, V, this.Nx, this.Nr);
  }
  dispose() { if (this._h) { addon.destroy(this._h); this._h = null; } }
}

class GPUSpray {
  constructor(solver) {
    this.s = solver;
    this.blendOk = true;   // fixed-point integer atomics on the device stand in for EXT_float_blend (spray.js:215)
    this.ready = false;
  }
  resize() { this.ready = true; this.reset(); }
  setInjector(inj) {
    if (!inj || !this.ready) return;
    addon.sprayInjector(this.s._h, inj.tab);
  }
  reset() { if (this.ready) addon.sprayReset(this.s._h); }
  setCount(n) { addon.sprayCount(this.s._h, n); }
  // the droplet overlay (spray.js:299-310) is visual; the render side is out of scope
  draw() {}
  stats(P) { return addon.sprayStats(this.s._h); }
}

// the C++ restatement of buildSimSetup (simsetup.js:7-126) for hosts without the browser modules
function buildSimSetup(cfg) {
  const s = addon.buildSimSetup(cfg);
  s.initialState = () => ({ A: new Float32Array(s.init.A), B: new Float32Array(s.init.B) });
  s.contour.handle = s.handle;
  return s;
}

// HUD read-outs (simsetup.js:352-384) on the C++ setup; same arguments as the reference
// (app.js:568 passes the setup, app.js:689 passes setup.contour)
function wallSeparation(setup, row, paSi) { return addon.wallSeparation(setup.handle, row, paSi); }
function summerfieldPoint(contour, Pc, g, pa) { return addon.summerfieldPoint(contour.handle, Pc, g, pa); }

module.exports = { GPUSolver, GPUSpray, buildSimSetup, wallSeparation, summerfieldPoint, deviceAvailable: addon.deviceAvailable };
