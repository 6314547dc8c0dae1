"""Autotuner and its per-shape cache file (spfft_amd/csrc/tuner.cpp).

Each case runs in a fresh process (the choices are per process), one at a
time: a batch of 4 FP32 transforms of 64^3 (whole-plane, persistent and
class-split kernels are candidates, the sphere's sticks are runs) and a
single FP64 transform. The first process times the candidates and writes the
cache file; the second reads the file, times nothing, takes the same choices
and produces the same bits. Outputs are checked against the oracle.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r'''
import sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import torch
import spfft_amd as sp
from oracle import oracle as orc
from tests.workloads import complex_values, sphere_triplets

out_path = sys.argv[2]
t = sphere_triplets(20)
dims = (64, 64, 64)
o = orc.OraclePlan(False, dims, t)
res = {}
# batch of 4 FP32 transforms
trs = [sp.TransformFloat(-1, sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, *dims, len(t),
                         sp.SPFFT_INDEX_TRIPLETS, t) for _ in range(4)]
vals = [complex_values(0x5EED0A00 + b, len(t)) for b in range(4)]
ins = [torch.from_numpy(v.astype(np.complex64)).cuda() for v in vals]
spaces = [torch.empty(dims[::-1], dtype=torch.complex64, device="cuda") for _ in range(4)]
outs = [torch.empty(len(t), dtype=torch.complex64, device="cuda") for _ in range(4)]
p = lambda x: torch.view_as_real(x).data_ptr()
sp.multi_transform_backward(trs, [p(x) for x in ins], [p(x) for x in spaces])
sp.multi_transform_forward(trs, [p(x) for x in spaces], [p(x) for x in outs],
                           [sp.SPFFT_FULL_SCALING] * 4)
torch.cuda.synchronize()
for b in range(4):
    want = o.backward(vals[b])
    assert orc.rel_l2(spaces[b].cpu().numpy().astype(np.complex128), want) <= 1e-5
    res[f"space{b}"] = spaces[b].cpu().numpy()
    res[f"out{b}"] = outs[b].cpu().numpy()
# one FP64 transform
tr = sp.Transform(-1, sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, *dims, len(t), sp.SPFFT_INDEX_TRIPLETS, t)
v = complex_values(0x5EED0A10, len(t))
din = torch.from_numpy(v).cuda()
space = torch.empty(dims[::-1], dtype=torch.complex128, device="cuda")
out = torch.empty(len(t), dtype=torch.complex128, device="cuda")
tr.backward(p(din), p(space))
tr.forward(p(space), p(out), sp.SPFFT_FULL_SCALING)
torch.cuda.synchronize()
want = o.backward(v)
assert orc.rel_l2(space.cpu().numpy(), want) <= 1e-12
assert orc.rel_l2(out.cpu().numpy(), o.forward(want, True)) <= 1e-12
res["single_space"] = space.cpu().numpy()
res["single_out"] = out.cpu().numpy()
np.savez(out_path, **res)
print("child done", flush=True)
'''


def run_child(tmp_path, tag, cache):
    env = dict(os.environ)
    env["SPFFT_AMD_TUNE_CACHE"] = str(cache)
    env["SPFFT_AMD_TUNE_LOG"] = "1"
    env.pop("SPFFT_AMD_PLANE", None)
    script = tmp_path / "child.py"
    script.write_text(CHILD)
    out = tmp_path / f"{tag}.npz"
    print(f"tuner child {tag}", flush=True)
    r = subprocess.run([sys.executable, "-u", str(script), ROOT, str(out)], env=env,
                       capture_output=True, text=True, timeout=240, cwd=ROOT)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stderr, dict(np.load(out))


def test_tuner_cache_round_trip(tmp_path):
    cache = tmp_path / "cache" / "tune.txt"
    log1, res1 = run_child(tmp_path, "first", cache)
    assert "spfft_amd tune:" in log1, log1
    text = cache.read_text()
    lines = [ln for ln in text.splitlines() if ln and not ln.startswith("#")]
    # batched backward and forward (x4) and the single transform's two directions
    assert any(":x4 " in ln and ":bwd:" in ln for ln in lines), text
    assert any(":x4 " in ln and ":fwd:" in ln for ln in lines), text
    assert any(":x1 " in ln and ":f64:" in ln for ln in lines), text
    for ln in lines:
        assert "plane=" in ln and "slots=" in ln and "zruns=" in ln, ln
    # an unreadable line is ignored, the rest still serves
    cache.write_text(text + "garbage line without fields\n")
    log2, res2 = run_child(tmp_path, "second", cache)
    assert "spfft_amd tune:" not in log2, log2
    for k in res1:
        np.testing.assert_array_equal(res1[k], res2[k], err_msg=k)


def test_tuner_cache_off(tmp_path):
    log, _ = run_child(tmp_path, "off", "off")
    assert "spfft_amd tune:" in log
    assert not any(p.name.startswith("tune") for p in tmp_path.rglob("*"))


HOST_FIRST = r'''
import sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import spfft_amd as sp
from oracle import oracle as orc
from tests.workloads import complex_values, sphere_triplets
# The first call of each shape takes pageable host buffers (staged through the
# grid's work buffer): the choices must not be timed on that buffer.
for n, prec in ((32, "float"), (64, "double"), (96, "float"), (192, "double")):
    dims = (n, n, n)
    t = sphere_triplets(0.3 * n)
    v = complex_values(0x5EED0A20 + n, len(t))
    cls = sp.Transform if prec == "double" else sp.TransformFloat
    rd = np.float64 if prec == "double" else np.float32
    tol = 1e-12 if prec == "double" else 1e-5
    tr = cls(-1, sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, *dims, len(t), sp.SPFFT_INDEX_TRIPLETS, t)
    space = np.zeros(2 * n ** 3, rd)
    tr.backward(np.ascontiguousarray(v.view(np.float64).astype(rd)), space)
    o = orc.OraclePlan(False, dims, t)
    want = o.backward(v)
    assert orc.rel_l2(space.astype(np.float64).view(np.complex128).reshape(n, n, n), want) <= tol
    out = np.zeros(2 * len(t), rd)
    tr.forward(np.ascontiguousarray(want.view(np.float64).ravel().astype(rd)), out,
               sp.SPFFT_FULL_SCALING)
    assert orc.rel_l2(out.astype(np.float64).view(np.complex128), o.forward(want, True)) <= tol
print("child done", flush=True)
'''


def test_first_call_with_host_pointers(tmp_path):
    env = dict(os.environ)
    env["SPFFT_AMD_TUNE_CACHE"] = "off"
    env.pop("SPFFT_AMD_PLANE", None)
    script = tmp_path / "host_first.py"
    script.write_text(HOST_FIRST)
    print("tuner child host_first", flush=True)
    r = subprocess.run([sys.executable, "-u", str(script), ROOT], env=env, capture_output=True,
                       text=True, timeout=240, cwd=ROOT)
    assert r.returncode == 0, r.stdout + r.stderr
