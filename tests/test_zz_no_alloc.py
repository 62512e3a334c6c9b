"""emu_sv_hpsi_static: H.psi that allocates nothing (the library's static
device scratch instead of a caller workspace), against emu_sv_hpsi on the
same inputs: two calls in a row, the second smaller, bit for bit (the same
kernels and reduction order), with the fused dots; sizes beyond the static
array are rejected."""

import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu


def _block(n, seed, phase):
    from emulators_amd import _native as native

    g = torch.Generator().manual_seed(seed)
    om = torch.rand(n, dtype=torch.float64, generator=g) * 8
    de = torch.randn(n, dtype=torch.float64, generator=g) * 4
    ph = torch.randn(n, dtype=torch.float64, generator=g) if phase else torch.zeros(n, dtype=torch.float64)
    u = torch.triu(torch.rand(n, n, dtype=torch.float64, generator=g) * 20, 1)
    block, cplx = native.pack_hamiltonian(om.numpy(), de.numpy(), ph.numpy(), (u + u.T).numpy())
    return torch.from_numpy(block).to("cuda"), cplx


def _static_call(n, params, sigma, v, w, dots, cplx, keep=False, kernel="default", tile_log=0,
                 min_contig=0, first_log=0):
    from emulators_amd import _native as native

    flags = (native.EMU_SV_FLAG_COMPLEX if cplx else 0) | native.kernel_flags(kernel, 0, first_log)
    if keep:
        flags |= native.EMU_SV_FLAG_KEEP_DIAG
    return native.LIB.emu_sv_hpsi_static(
        n, native.dtype_code(v.dtype), ctypes.c_void_p(params.data_ptr()), sigma.real, sigma.imag,
        ctypes.c_void_p(v.data_ptr()), 1.0, None, 1.0, ctypes.c_void_p(w.data_ptr()),
        ctypes.c_void_p(dots.data_ptr()) if dots is not None else None, tile_log, min_contig, flags,
        native._stream(v.device))


@pytest.mark.parametrize("kernel,tile_log,min_contig,first_log",
                         [("default", 0, 0, 0), ("tile", 12, 4, 11), ("lds", 11, 2, 0)])
def test_static_scratch_hpsi_matches_workspace_hpsi(gpu_device, kernel, tile_log, min_contig,
                                                    first_log):
    from emulators_amd import _native as native

    sigma = -0.01j
    kw = dict(kernel=kernel, tile_log=tile_log, min_contig=min_contig, first_log=first_log)
    for n, dtype, seed, phase in ((20, torch.complex128, 1, True), (14, torch.complex64, 2, False)):
        params, cplx = _block(n, seed, phase)
        g = torch.Generator(device=gpu_device).manual_seed(seed)
        v = torch.randn(2**n, dtype=dtype, device=gpu_device, generator=g)
        ref = torch.empty_like(v)
        ref_dots = native.hpsi(n, params, sigma, v, 1.0, ref, dots=True, complex_coef=cplx,
                               **kw).clone()
        for keep in (False, True):  # the second call may keep the stamped diagonal table
            w = torch.full_like(v, float("nan"))
            dots = torch.full((7,), float("nan"), dtype=torch.float64, device=gpu_device)
            rc = _static_call(n, params, sigma, v, w, dots, cplx, keep=keep, **kw)
            assert rc == native.EMU_SV_OK, native.LIB.emu_sv_last_error().decode()
            torch.cuda.synchronize()
            assert torch.equal(w, ref), (n, dtype, keep)
            assert torch.equal(dots, ref_dots), (n, dtype, keep, dots, ref_dots)


def test_static_scratch_rejects_sizes_beyond_the_array(gpu_device):
    from emulators_amd import _native as native

    params, cplx = _block(25, 3, False)
    v = torch.zeros(2, dtype=torch.complex128, device=gpu_device)  # never read: rejected first
    rc = _static_call(25, params, 1.0 + 0j, v, v.clone(), None, cplx)
    assert rc == native.EMU_SV_EINVAL
    assert "static array" in native.LIB.emu_sv_last_error().decode()
