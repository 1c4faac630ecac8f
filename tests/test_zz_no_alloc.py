"""hs_metropolis_run_static, the allocation-free variant of hs_metropolis_run (its scratch is a
file-scope device array), called through the C-ABI: two calls in a row, the second on a smaller
instance, each bit-identical to hs_metropolis_run with a workspace on a twin instance (spins,
fields and their compensations, time and counters); a call that needs more scratch than the
array holds is refused before anything is launched."""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _device_line():
    from hotspice_amd import _lib
    lib = _lib.load()
    cnt, name = C.c_int32(0), C.create_string_buffer(128)
    rc = lib.hs_device_info(C.byref(cnt), name, 128)
    return f"hs_device_info rc={rc} count={cnt.value} name={name.value!r} last_error={lib.hs_last_error()!r}"


def _twins(n, seed, r):
    """Two identical OOP square instances (dipolar window + exchange, grid selection at r)."""
    import hotspice_amd as hs
    from hotspice_amd.workloads import edge_cutoff, kB
    out = []
    for _ in range(2):
        rng = np.random.default_rng(seed)
        kT = kB * 600
        out.append(hs.ASI.OOP_Square(1e-6, n, T=600, E_B=20 * kT, moment=1.1e-15,
                                     pattern=np.where(rng.random((n, n)) < .5, -1., 1.), seed=seed,
                                     energies=[hs.DipolarEnergy(cutoff_range=edge_cutoff(3e-6)),
                                               hs.ExchangeEnergy(J=0.1 * kT)],
                                     params=hs.SimParams(UPDATE_SCHEME='Metropolis', REDUCED_KERNEL_SIZE=3,
                                                         SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=0)))
    sel = out[0]._select_geometry(r)
    return out[0], out[1], sel, out[1]._select_geometry(r)


def _run(steps):
    from hotspice_amd import _lib
    return _lib.Run(t_max=math.inf, mc_max=math.inf, attempt_freq=1e10, max_steps=steps, log=0, nblocks=0, begin=1,
                    first_free=0)


def _snapshot(mm):
    d = mm._dev
    st = d.read_state()
    return dict(m=d.m.cpu().numpy(), E=d.E.cpu().numpy(), Ec=d.Ec.cpu().numpy(),
                counters=(st.t, st.attempts, st.switches, st.step, st.steps_done, st.done, st.err))


@pytest.mark.parametrize("sizes", [((200, 1.0e-6), (96, 2.5e-6), (40, 3e-6))])
def test_static_scratch_equals_workspace(sizes):
    """200 x 200 at r = 1 cell (presampled steps, >1024 samples), then 96 x 96 (one selection
    workgroup per step), then 40 x 40 (the persistent one-workgroup kernel): each call follows a
    larger one on the same static scratch."""
    from hotspice_amd import _lib
    lib = _lib.load()
    print(_device_line(), flush=True)
    for n, r in sizes:
        a, b, sel_a, sel_b = _twins(n, 7 + n, r)
        need = int(lib.hs_metropolis_ws_bytes(C.byref(b._dev.inst), C.byref(sel_b)))
        assert 0 < need <= _lib.HS_METRO_STATIC_SCRATCH_BYTES
        print(f"static scratch: {n}x{n}, nsc {sel_b.nsc}, {need} bytes", flush=True)
        for m in (a, b):
            m._pre_step()
        a._dev.metropolis_run(sel_a, _run(12))
        d = b._dev
        _lib.check(lib.hs_metropolis_run_static(C.byref(d.inst), C.byref(sel_b), C.byref(_run(12)), d.ptr(d.state),
                                                None, None, d.stream()))
        for m in (a, b):
            m._after_device_step()
        sa, sb = _snapshot(a), _snapshot(b)
        for k in ("m", "E", "Ec"):
            np.testing.assert_array_equal(sa[k], sb[k], err_msg=f"{n}: {k}")
        assert sa["counters"] == sb["counters"], n
        assert sa["counters"][2] > 0 and sa["counters"][-1] == 0


def test_static_scratch_refuses_oversize():
    """A selection of 2^20 samples needs more than the array: HS_ERR_ARG and a message, no launch."""
    from hotspice_amd import _lib
    lib = _lib.load()
    a, _, sel, _ = _twins(40, 3, 3e-6)
    big = _lib.Select.from_buffer_copy(sel)
    big.scheme, big.nsc = 2, 1 << 20
    assert int(lib.hs_metropolis_ws_bytes(C.byref(a._dev.inst), C.byref(big))) > _lib.HS_METRO_STATIC_SCRATCH_BYTES
    before = _snapshot(a)
    d = a._dev
    rc = lib.hs_metropolis_run_static(C.byref(d.inst), C.byref(big), C.byref(_run(4)), d.ptr(d.state), None, None,
                                      d.stream())
    assert rc == -1
    assert b"static scratch" in lib.hs_last_error()
    after = _snapshot(a)
    np.testing.assert_array_equal(before["m"], after["m"])
    assert before["counters"] == after["counters"]
