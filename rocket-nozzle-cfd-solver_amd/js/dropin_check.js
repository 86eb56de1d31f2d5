// dropin_check.js — replays the reference app's call sequence against the drop-in:
// restart() (app.js:229-252), frame() stepping with the fenced asynchronous
// read-back protocol (app.js:576-603), the synchronous readback() (app.js:534-539)
// and the self-test's speed loop (app.js:1295-1311). Prints one JSON line.
//
// usage: node dropin_check.js '<cfg JSON>' <frames> <stepsPerFrame> [parcels]
'use strict';
const { GPUSolver, GPUSpray, buildSimSetup, wallSeparation, summerfieldPoint } = require('./gpu_solver.js');

const cfg = JSON.parse(process.argv[2]);
const frames = parseInt(process.argv[3] || '5', 10);
const perFrame = parseInt(process.argv[4] || '10', 10);
const parcels = parseInt(process.argv[5] || '0', 10);

// restart(): build the setup, size the solver, upload geometry / injector / state
const setup = buildSimSetup(cfg);
const solver = new GPUSolver(null);
solver.resize(setup.Nx, setup.Nr);
solver.setGeometry(setup.geo);
if (!solver.spray) { solver.spray = new GPUSpray(solver); solver.spray.resize(); }
if (parcels) solver.spray.setCount(parcels);
solver.spray.setInjector(setup.injector);
const init = setup.initialState();
solver.setState(init.A, init.B);
solver.setProbes(setup);
const P = setup.P;

// frame(): step, then the non-blocking read-back of app.js:588-601
let polled = 0, requested = 0, lastAsync = null;
for (let f = 0; f < frames; f++) {
  solver.step(perFrame, P);
  const res = solver.pollAsync();
  if (res) { polled++; lastAsync = res; }
  if (!solver._async) {
    const st = solver.state;
    const reqs = [
      { name: 'step', tex: solver.stepInfoTex, x: 0, y: 0, w: 1, h: 1 },
      { name: 'colA', tex: st.A, x: setup.iExit, y: 0, w: 1, h: setup.Nr },
      { name: 'colB', tex: st.B, x: setup.iExit, y: 0, w: 1, h: setup.Nr },
      { name: 'chA', tex: st.A, x: setup.iChamber, y: 0, w: 1, h: 1 },
      { name: 'wall', tex: solver.drawWallRow(), x: 0, y: 0, w: setup.Nx, h: 1 },
    ];
    if (solver.requestAsync(reqs, f)) requested++;
  }
}
while (solver._async) { const r = solver.pollAsync(); if (r) { polled++; lastAsync = r; } }

// the self-test speed loop: 10 x step(10) closed by a synchronous step-info read
const t0 = process.hrtime.bigint();
for (let k = 0; k < 10; k++) solver.step(10, P);
solver.readStepInfo();
const ms = Number(process.hrtime.bigint() - t0) / 1e6 / 100;

// readback(): synchronous step info, exit column and chamber cell
const si = solver.readStepInfo();
const col = solver.readColumn(setup.iExit);
const ch = solver.readCell(setup.iChamber, 0);
const all = solver.readAll();
let sumA = 0, sumB = 0, finite = true;
for (let k = 0; k < all.A.length; k++) { sumA += all.A[k]; sumB += all.B[k]; if (!isFinite(all.A[k])) finite = false; }
// exitPerformance (simsetup.js:314-337) on the read-back column, with the ambient pressure
// at the double precision of the reference's P (the setup's float uniforms round it)
const pa = setup.probes ? setup.probes.pa : P.amb[3];
let F = 0;
for (let j = 0; j < setup.jExit; j++) {
  const k = j * 4, rho = Math.max(col.A[k], 1e-9), u = col.A[k + 1] / rho, v = col.A[k + 2] / rho;
  const G = col.B[k + 2];
  const p = Math.max((col.A[k + 3] - 0.5 * rho * (u * u + v * v)) / G, 1e-9);
  const dA = 2 * Math.PI * Math.abs(setup.geo.Ax[j * setup.geo.W1 + setup.iExit + 1]) * setup.dx * setup.dx;
  F += (rho * u * u + p - pa) * dA;
}
const integ = solver.integrals(P);
// render.js derive(V) for the Mach field (app.js V() fills the view), and the HUD
// read-outs of app.js processWall (568) and the separation row (689)
const V = { fieldIndex: 1, range: [0, 4], RaN: P.gas[0], Rs: P.Rs, chemK: P.chemK, Tc: 300, Ta: 300, schlGain: 1, dx: setup.dx };
const der = solver.derive(V);
let sumDerived = 0;
for (let k = 0; k < der.length; k++) sumDerived += der[k];
const wallRow = solver._readRect(solver.drawWallRow(), 0, 0, setup.Nx, 1);
const sep = wallSeparation(setup, wallRow, P.amb[3] * 101325);
const sfp = summerfieldPoint(setup.contour, 10e5, 1.4, 1e5);

const out = {
  Nx: setup.Nx, Nr: setup.Nr, stepCount: solver.stepCount, stepInfo: Array.from(si),
  chamberA: Array.from(ch.A), sumA, sumB, finite, thrustJS: F * 101325, thrustDevice: integ.F,
  requested, polled, lastTag: lastAsync ? lastAsync.tag : null,
  spray: solver.spray.stats(P), msPerStep: ms, sumDerived, wallRow: Array.from(wallRow), sep, summerfield: sfp,
};
solver.dispose();
console.log(JSON.stringify(out));
