"""The frame readout (pack, kernels.wgsl.js:1346-1355, read through frame(tEnd, K, readOut),
solver-gpu.js:607-622): per interior cell, row-major, the depth and the speed sqrt(hu^2 + hv^2) / h where
h > hdry (0 elsewhere), as FP32 pairs.

Checked, through the Python binding and through the JavaScript class under Node: the depths equal
readState().h bit for bit (both are the FP32 state), the speeds equal hypot(hu, hv) / h of readState's
FP32 momenta within FP32 rounding (rtol 1e-6: the device may fuse the sum of squares), and the depths are
within 1 mm of the FP64 oracle on the same inputs; frame(.., false) returns no readout (None / null) and
advances the same steps.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from oracle import refcases as rc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
TOL_H = 1e-3


def _case():
    sc = rc.cloudburst(80)
    opts = dict(nx=sc["nx"], ny=sc["ny"], dx=sc["dx"], z=sc["z"], mask=sc["mask"], manning=sc["manning"],
                bc=sc["bc"], infil=sc["infil"])
    return sc, opts


def _check_pack(out, h, hu, hv, hdry):
    out = np.asarray(out, dtype=np.float32)
    assert out.shape == (2 * h.size,)
    d, sp = out[0::2], out[1::2]
    h32, hu32, hv32 = (np.asarray(a, dtype=np.float64).astype(np.float32) for a in (h, hu, hv))
    np.testing.assert_array_equal(d, h32)
    wet = h32 > np.float32(hdry)
    want = np.sqrt(hu32[wet].astype(np.float64) ** 2 + hv32[wet].astype(np.float64) ** 2) / h32[wet]
    np.testing.assert_allclose(sp[wet], want, rtol=1e-6, atol=0)
    assert np.all(sp[~wet] == 0)
    return int(wet.sum()), float(sp.max())


def test_frame_readout_python():
    import floodsim_amd as fsa
    sc, opts = _case()
    g = fsa.GPUSolver(**opts)
    g2 = fsa.GPUSolver(**opts)
    c = oracle.CPUSolver(**opts, hdry=g.hdry)
    for o in (g, g2, c):
        o.setDepth(sc["depth"])
        o.rainRate = 150 * rc.MMHR
    out = None
    for t in range(30, 301, 30):
        out = g.frame(float(t), 1 << 20, True)
        assert g2.frame(float(t), 1 << 20, False) is None
        c.run(float(t))
    s = g.readState()
    wet, vmax = _check_pack(out, s["h"], s["hu"], s["hv"], g.hdry)
    d = float(np.max(np.abs(out[0::2].astype(np.float64) - c.field("h"))))
    print(f"\nframe readout: {wet} wet cells, max speed {vmax:.3f} m/s, max|dh| vs FP64 {d:.2e} m, "
          f"steps {g.steps}/{c.steps}")
    assert wet > 0 and vmax > 0
    assert d <= TOL_H
    assert g2.steps == g.steps
    np.testing.assert_array_equal(g2.readState()["h"], s["h"])


JS = r"""
const path = require('path');
const { GPUSolver } = require(path.join(process.cwd(), 'floodsim_amd', 'js', 'solver-hip.js'));
const c = JSON.parse(require('fs').readFileSync(process.argv[1], 'utf8'));
(async () => {
  const s = new GPUSolver(null, { nx: c.nx, ny: c.ny, dx: c.dx, z: Float64Array.from(c.z), manning: Float64Array.from(c.manning),
    mask: Uint8Array.from(c.mask), bc: c.bc, infil: Float64Array.from(c.infil), hazard: false });
  await s.ready;
  s.setDepth((x, y) => c.depth[Math.floor(y / c.dx) * c.nx + Math.floor(x / c.dx)]);
  s.rainRate = c.rain;
  let out = null, none = 0;
  for (let t = 30; t <= 300; t += 30) out = await s.frame(t, 1 << 20, true);
  if ((await s.frame(330, 4, false)) === null) none = 1;
  const st = await s.readState();
  out = await s.frame(330, 0, true);
  const st2 = await s.readState();
  process.stdout.write(JSON.stringify({ out: Array.from(out), h: Array.from(st2.h), hu: Array.from(st2.hu),
    hv: Array.from(st2.hv), steps: s.steps, none, hdry: s.hdry }) + '\n');
  s.destroy();
})().catch((e) => { process.stderr.write(String(e && e.stack || e) + '\n'); process.exit(1); });
"""


@pytest.mark.skipif(NODE is None, reason="node is not installed")
def test_frame_readout_js(tmp_path):
    sc, opts = _case()
    X, Y = np.meshgrid((np.arange(sc["nx"]) + 0.5) * sc["dx"], (np.arange(sc["ny"]) + 0.5) * sc["dx"])
    depth = np.asarray(sc["depth"](X, Y, np.asarray(sc["z"]).reshape(sc["ny"], sc["nx"])), dtype=np.float64)
    depth = np.broadcast_to(depth, X.shape).reshape(-1)
    case = dict(nx=sc["nx"], ny=sc["ny"], dx=sc["dx"], z=list(map(float, np.asarray(sc["z"]).reshape(-1))),
                manning=list(map(float, np.broadcast_to(np.asarray(sc["manning"], dtype=float), (sc["nx"] * sc["ny"],)))),
                mask=list(map(int, np.asarray(sc["mask"]).reshape(-1))), bc=sc["bc"],
                infil=list(map(float, np.broadcast_to(np.asarray(sc["infil"], dtype=float), (sc["nx"] * sc["ny"],)))),
                depth=list(map(float, depth)), rain=150 * rc.MMHR)
    f = tmp_path / "case.json"
    f.write_text(json.dumps(case))
    print("\nnode frame readout case", flush=True)
    r = subprocess.run([NODE, "-e", JS, str(f)], cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    o = json.loads(r.stdout)
    assert o["none"] == 1
    wet, vmax = _check_pack(o["out"], np.array(o["h"]), np.array(o["hu"]), np.array(o["hv"]), o["hdry"])
    print(f"JS frame readout: {wet} wet cells, max speed {vmax:.3f} m/s, {o['steps']} steps")
    assert wet > 0
