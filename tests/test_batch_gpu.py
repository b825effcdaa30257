"""Cross-mesh batched assembly (emg_assemble_iso_batched): geometry variants of one
mesh topology assembled in one fused launch pair equal, bit for bit, the single-mesh
assembly of each variant (itself pinned against the oracle in test_assembly_gpu.py)."""
import numpy as np
import pytest

from emerge_amd import meshgen
from emerge_amd.assembly import EXACT_FACE_MASS, REFERENCE_GEOMETRY, GPUPattern

pytestmark = pytest.mark.gpu


def _variants(mesh, M, seed=0):
    """M coordinate sets: the mesh stretched along x, y, z by factors in [0.9, 1.1] (positive
    scalings keep every tet's orientation), and per-tet eps_r = er (1 - j tan d)."""
    rng = np.random.default_rng(seed)
    nodes = np.empty((M, mesh.n_nodes, 3))
    for b in range(M):
        nodes[b] = mesh.nodes.T * rng.uniform(0.9, 1.1, 3)[None, :]
    er = rng.uniform(1, 10, (M, mesh.n_tets))
    eps = er * (1 - 1j * rng.uniform(0, 0.02, (M, mesh.n_tets)))
    nu = np.ones((M, mesh.n_tets), dtype=np.complex128)
    return nodes, eps, nu


@pytest.mark.parametrize("flags", [0, EXACT_FACE_MASS, REFERENCE_GEOMETRY])
def test_batched_assembly_equals_single(cuda, flags):
    import torch
    mesh = meshgen.waveguide_wr90()
    pat = GPUPattern(mesh)
    M = 6
    nodes, eps, nu = _variants(mesh, M)
    E, B = pat.assemble_iso_batched(nodes, eps, nu, flags)
    base = pat.nodes
    try:
        for b in range(M):
            pat.nodes = torch.from_numpy(np.ascontiguousarray(nodes[b])).to(pat.device)
            E1 = torch.empty(pat.nnz, dtype=torch.complex128, device=pat.device)
            B1 = torch.empty_like(E1)
            pat.assemble_iso_device(torch.from_numpy(eps[b]).to(pat.device), torch.from_numpy(nu[b]).to(pat.device),
                                    flags, E1, B1)
            assert torch.equal(E[b], E1) and torch.equal(B[b], B1), b
    finally:
        pat.nodes = base
    # the variants differ: the batch did not assemble one mesh M times
    assert not torch.equal(E[0], E[1])


@pytest.mark.parametrize("hexes,M", [((6, 3, 8), 128), ((15, 7, 32), 32)])
def test_batched_assembly_against_one_by_one(cuda, hexes, M):
    """M WR-90 variants: one batched launch pair against M single ones.  Small meshes (864
    tets, whose own launch cannot fill the GPU) must gain; 20,160-tet meshes, which fill it alone,
    must not lose more than timing noise."""
    import torch
    mesh = meshgen.waveguide_wr90(*hexes)
    pat = GPUPattern(mesh)
    nodes, eps, nu = _variants(mesh, M, seed=1)
    dn = torch.from_numpy(nodes).to(pat.device)
    de = torch.from_numpy(eps).to(pat.device)
    dnu = torch.from_numpy(nu).to(pat.device)
    E1 = torch.empty(pat.nnz, dtype=torch.complex128, device=pat.device)
    B1 = torch.empty_like(E1)
    singles = [dn[b].contiguous() for b in range(M)]

    def loop():
        base = pat.nodes
        for b in range(M):
            pat.nodes = singles[b]
            pat.assemble_iso_device(de[b], dnu[b], 0, E1, B1)
        pat.nodes = base

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        a, z = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        z.record()
        torch.cuda.synchronize()
        return a.elapsed_time(z) / reps

    t_batch = timed(lambda: pat.assemble_iso_batched(dn, de, dnu))
    t_loop = timed(loop)
    mt = M * mesh.n_tets / 1e6
    print(f"{M} x {mesh.n_tets} tets: batched {t_batch:.3f} ms ({mt / t_batch * 1e3:.0f} Mtet/s), "
          f"one by one {t_loop:.3f} ms ({mt / t_loop * 1e3:.0f} Mtet/s)", flush=True)
    if mesh.n_tets < 5000:
        assert t_batch < 0.5 * t_loop          # measured 0.38 against 1.97 ms
    else:
        assert t_batch < 1.5 * t_loop          # measured 2.24 against 2.18 ms: no loss beyond timing noise
