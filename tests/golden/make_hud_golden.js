// make_hud_golden.js — input/output vectors of the reference's HUD read-outs
// wallSeparation (simsetup.js:352-373) and summerfieldPoint (simsetup.js:375-384)
// on synthetic near-wall rows, for a few setups.
//
// Runs in the development container only, under node, on the reference's own
// source files read from the directory given as argv[2]. Only the numeric inputs
// and results are written out.
//
// usage: node make_hud_golden.js <reference source dir> <output json>
'use strict';
const fs = require('fs');
const path = require('path');
const vm = require('vm');

const srcDir = process.argv[2];
const outFile = process.argv[3];
const files = ['physics.js', 'cutcell.js', 'simsetup.js'];
let code = files.map((f) => fs.readFileSync(path.join(srcDir, f), 'utf8')).join('\n;\n');
code += '\n;this.__api = { buildSimSetup, wallSeparation, summerfieldPoint };\n';
const ctx = vm.createContext({ Math, Float32Array, Float64Array, Uint8Array, Int32Array, Array, Object, JSON, Infinity, NaN, isFinite });
vm.runInContext(code, ctx, { filename: 'reference-host.js' });
const api = ctx.__api;

function baseCfg() {
  return {
    Dt: 0.023, De: 0.119, type: 'bell', bellPct: 80, coneAngle: 15, CR: 6, chamberLen: 1.6, Pc: 196e5, Tc: 3670,
    gamma: 1.22, Mw: 22, alt: 0, mach: 0, base: false, bodyRatio: 1.5, domainL: 1.2, Nx: 256, aspect: 0.28,
    startMode: 'ramp', rampMs: 0.12, cfl: 0.6, order2: true, wallModel: 'cut', minCells: 8, viscous: true,
    afterburn: true, qAb: 6.8, sAir: 1.8, Tw: 800, inletAbsorb: 1, prop: 'kerolox', injMode: false, injType: 'shower',
    injRings: 6, injDP: 20, injD: 1.5, injTMR: 1, injAngle: 30, OF: 2.3, film: 0, baffles: 0, baffleH: 30, cavD: 0, cavW: 40,
  };
}
const cases = {
  demo_256: baseCfg(),
  overexp_192: Object.assign(baseCfg(), { prop: 'hydrolox', Tc: 3400, gamma: 1.2, Mw: 13.5, Pc: 120e5, Dt: 0.020, De: 0.160, CR: 5, chamberLen: 1.4, domainL: 1.4, Nx: 192 }),
  cone_160: Object.assign(baseCfg(), { type: 'cone', coneAngle: 15, Dt: 0.020, De: 0.060, Pc: 30e5, domainL: 0.6, Nx: 160 }),
};

// deterministic generator for the synthetic rows
let seed = 12345;
function rnd() { seed = (seed * 1103515245 + 12345) % 2147483648; return seed / 2147483648; }

// near-wall rows (u, v, p, c per column, normalised) of a few kinds: attached, separated
// near the exit with supersonic flow ahead, separated with subsonic flow ahead, slow
// film with gaps, and random mixtures
function rows(s) {
  const out = [];
  const Nx = s.Nx, iT = s.iThroat, iE = s.iExit;
  const kinds = ['attached', 'separated', 'subsonic', 'gappy', 'random', 'random', 'random', 'separated'];
  for (const kind of kinds) {
    const row = new Float32Array(Nx * 4);
    const iS = Math.round(iT + (iE - iT) * (0.4 + 0.5 * rnd()));
    for (let i = 0; i < Nx; i++) {
      let u = 2.5 + 0.5 * rnd(), v = 0.3 * rnd(), p = 0.2 + 0.3 * rnd(), c = 1.0 + 0.2 * rnd();
      if (kind === 'separated' && i >= iS) { u = -0.3 * rnd(); p = 0.9 + 0.2 * rnd(); }
      if (kind === 'subsonic') { u = (i >= iS ? -0.2 : 0.5) * rnd(); }
      if (kind === 'gappy' && i >= iS && (i % 3 === 0)) { u = -0.1; }
      if (kind === 'random' && rnd() < 0.3) { u = -rnd(); }
      row[i * 4] = u; row[i * 4 + 1] = v; row[i * 4 + 2] = p; row[i * 4 + 3] = c;
    }
    out.push({ kind, row });
  }
  return out;
}

const result = {};
for (const [name, cfg] of Object.entries(cases)) {
  const s = api.buildSimSetup(JSON.parse(JSON.stringify(cfg)));
  const rs = rows(s).map(({ kind, row }) => {
    const pa = 101325 * (0.2 + rnd());
    const sep = api.wallSeparation(s, row, pa);
    return { kind, pa, row: Array.from(row), sep };
  });
  const sf = [];
  for (const pa of [101325, 50000, 20000, 5000, 1000]) {
    for (const Pc of [0.5 * cfg.Pc, cfg.Pc, 2 * cfg.Pc]) sf.push({ Pc, g: cfg.gamma, pa, frac: api.summerfieldPoint(s.contour, Pc, cfg.gamma, pa) });
  }
  result[name] = { cfg, rows: rs, summerfield: sf };
}
fs.writeFileSync(outFile, JSON.stringify(result));
console.log('wrote', Object.keys(result).length, 'HUD cases to', outFile);
