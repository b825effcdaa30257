"""emg_sweep_residual_noalloc (partial sums in the library's static device
array) against emg_sweep_residual (caller workspace): identical norms and
residual vectors on two calls in a row, the second on a smaller input."""
import numpy as np
import pytest

from emerge_amd import _lib, meshgen
from emerge_amd.assembly import GPUPattern

pytestmark = pytest.mark.gpu

def _noalloc(lib):
    return lib.emg_sweep_residual_noalloc


def _case(cuda, nx, ny, nz, nvec, seed):
    import torch
    mesh = meshgen.structured_box(nx, ny, nz, size=(1.0, 0.8, 0.6), jitter=0.1, seed=seed)
    rng = np.random.default_rng(seed)
    eps = rng.uniform(1, 10, mesh.n_tets) * (1 - 1j * rng.uniform(0, 0.02, mesh.n_tets))
    from oracle import oracle as O
    er, ur = O.material_tensors(eps)
    pat = GPUPattern(mesh, device=cuda)
    E, B = pat.assemble(er, ur)
    n = pat.n
    mask_h = (rng.uniform(size=n) > 0.1).astype(np.uint8)
    X = rng.standard_normal((n, nvec)) + 1j * rng.standard_normal((n, nvec))
    X *= mask_h[:, None]
    nq = 3
    Q = rng.standard_normal((nq, n)) + 1j * rng.standard_normal((nq, n))
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).astype(dt)).to(cuda)
    return dict(n=n, pat=pat, E=E, B=B, mask=t(mask_h, np.uint8), nvec=nvec,
                s=t(rng.uniform(1e3, 1e4, nvec), np.float64), X=t(X.ravel(), np.complex128),
                Q=t(Q, np.complex128), qidx=t(rng.integers(0, nq, nvec), np.int32),
                c=t(rng.standard_normal(nvec) + 1j * rng.standard_normal(nvec), np.complex128))


def _run(lib, d, noalloc: bool):
    import torch
    p = _lib.ptr
    nvec, n = d["nvec"], d["n"]
    R = torch.empty(n * nvec, dtype=torch.complex128, device=d["E"].device)
    norms = torch.full((2 * nvec,), float("nan"), dtype=torch.float64, device=d["E"].device)
    args = [n, p(d["pat"].rowptr), p(d["pat"].cols), p(d["E"]), p(d["B"]), p(d["mask"]), nvec, p(d["s"]),
            p(d["X"]), p(d["Q"]), p(d["qidx"]), p(d["c"]), None, None, p(R), p(norms)]
    if noalloc:
        _lib.check(_noalloc(lib)(*args, _lib.stream_ptr()), "emg_sweep_residual_noalloc")
    else:
        work = torch.empty(4096 * 16 * 2, dtype=torch.float64, device=d["E"].device)
        _lib.check(lib.emg_sweep_residual(*args, p(work), _lib.stream_ptr()), "emg_sweep_residual")
    torch.cuda.synchronize()
    return norms.cpu().numpy(), R.cpu().numpy()


def test_noalloc_matches_workspace_entry(cuda):
    lib = _lib.load()
    for shape, nvec, seed in (((8, 6, 5), 16, 11), ((4, 3, 3), 4, 12)):   # second call: smaller input
        print(f"no-alloc residual: box {shape}, nvec {nvec}", flush=True)
        d = _case(cuda, *shape, nvec, seed)
        n_ref, r_ref = _run(lib, d, noalloc=False)
        n_na, r_na = _run(lib, d, noalloc=True)
        assert np.all(np.isfinite(n_na))
        assert np.array_equal(n_na, n_ref), (n_na, n_ref)
        assert np.array_equal(r_na, r_ref)


def test_noalloc_rejects_bad_sizes(cuda):
    lib = _lib.load()
    rc = _noalloc(lib)(0, None, None, None, None, None, 3, None, None, None, None, None, None, None, None, None,
                       None)
    assert rc != 0
    assert b"nvec" in lib.emg_last_error()
