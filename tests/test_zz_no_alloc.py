"""nrt_setup_noalloc (device arrays in and out, nothing allocated, per-block counts
in a fixed per-device scratch array) against nrt_setup (host arrays) through the C
ABI: every output bit-identical, three calls in a row on the same device, each
smaller than the one before (the scratch keeps the previous call's contents), the
last one without HH elements. The test's own device buffers come from hipMalloc of
the HIP runtime the library itself links (test side only: the entry point under
test allocates nothing)."""
import ctypes as C

import numpy as np
import pytest

import neurite_hip as nh
import workloads as wl
from common import geometry_from_golden, load_golden

pytestmark = pytest.mark.gpu

N_FIELDS = ["dx", "c_m", "r_a", "r_m", "W", "K", "crit_dT", "rest"]
H_FIELDS = ["hh_m", "hh_h", "hh_n", "E_Na", "E_K", "E_L", "g_Na", "g_K", "G_L", "ls_Na", "ls_K",
            "hh_rest"]


def _host_setup(geom):
    H = nh.count_hh(geom)
    s, bufs = nh.setup_out_buffers(geom.N, H)
    g = geom.struct()
    assert nh.lib().nrt_setup(C.byref(g), 0, C.byref(s)) == 0, nh.lib().nrt_last_error(None)
    return H, s.dT_crit, bufs


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libneurite_hip links (loaded by
    its soname after the library, so the same copy)."""

    def __init__(self):
        nh.lib()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.held = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.held.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, n, dtype):
        out = np.empty(n, dtype)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.held:
            self.L.hipFree(p)
        self.held = []


def _device_setup(geom, hip, capacity):
    mats = (nh.Material * len(geom.materials))(*geom.materials)
    g = nh.GeometryStruct()
    g.num_elem = geom.N
    g.diameter = C.cast(hip.put(geom.diameter), C.POINTER(C.c_double))
    g.dx0 = C.cast(hip.put(geom.dx0), C.POINTER(C.c_double))
    g.epsilon = C.cast(hip.put(geom.epsilon), C.POINTER(C.c_double))
    g.material = C.cast(hip.put(geom.material), C.POINTER(C.c_int32))
    g.num_materials = len(geom.materials)
    g.materials = C.cast(hip.put(np.frombuffer(bytes(mats), np.uint8)), C.POINTER(nh.Material))
    g.myelinate = geom.myelinate
    g.E_MAX = geom.E_MAX
    cap = max(capacity, 1)
    outs = {k: (hip.put(np.full(geom.N, -7.0)), geom.N, np.float64) for k in N_FIELDS}
    outs.update({k: (hip.put(np.full(cap, -7.0)), cap, np.float64) for k in H_FIELDS})
    outs["hh_elem"] = (hip.put(np.full(cap, -7, np.int32)), cap, np.int32)
    s = nh.SetupOutStruct()
    for k, (ptr, _, dt) in outs.items():
        setattr(s, k, C.cast(ptr, C.POINTER(C.c_int32 if dt == np.int32 else C.c_double)))
    dT = hip.put(np.array([-1.0]))
    nhh = hip.put(np.array([-5], np.int32))
    L = nh.lib()
    L.nrt_setup_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int32,
                                    C.c_void_p, C.c_void_p]
    st = L.nrt_setup_noalloc(C.addressof(g), 0, None, C.addressof(s), capacity, dT, nhh)
    assert st == 0, L.nrt_last_error(None)
    res = (int(hip.get(nhh, 1, np.int32)[0]), float(hip.get(dT, 1, np.float64)[0]),
           {k: hip.get(ptr, n, dt) for k, (ptr, n, dt) in outs.items()})
    hip.free()
    return res


def _geometries():
    big, _ = wl.myelinated_axon()                                   # N 20,202, 200 HH
    mid = geometry_from_golden(load_golden("ex7_hh_axon_strain"))   # N 525, 25 HH, strained
    small = geometry_from_golden(load_golden("ex5_ct_y"))           # N 150, passive
    return [big, mid, small]


def test_setup_noalloc_matches_setup(nrt):
    hip = _Hip()
    geoms = _geometries()
    assert geoms[0].N > geoms[1].N > geoms[2].N
    for geom in geoms:
        print(f"[gpu-case] setup_noalloc N={geom.N}", flush=True)
        H, dT_ref, ref = _host_setup(geom)
        nhh, dT, got = _device_setup(geom, hip, H)
        assert nhh == H
        assert dT == dT_ref and dT > 0
        for k in N_FIELDS:
            assert np.array_equal(got[k], ref[k]), k
        for k in H_FIELDS + ["hh_elem"]:
            assert np.array_equal(got[k][:H], ref[k][:H]), k


def test_setup_noalloc_capacity_and_limits(nrt):
    """A capacity below the HH count fills only that many HH entries and reports the
    true count; a geometry above the scratch's element limit is refused."""
    hip = _Hip()
    geom = _geometries()[1]
    H, _, ref = _host_setup(geom)
    nhh, _, got = _device_setup(geom, hip, 10)
    assert nhh == H > 10
    assert np.array_equal(got["hh_elem"][:10], ref["hh_elem"][:10])
    assert np.array_equal(got["E_L"][:10], ref["E_L"][:10])
    g = nh.GeometryStruct()
    g.num_elem = (1 << 26) + 1
    s = nh.SetupOutStruct()
    d = C.c_double(0)
    n = C.c_int32(0)
    L = nh.lib()
    st = L.nrt_setup_noalloc(C.addressof(g), 0, None, C.addressof(s), 0, C.addressof(d),
                             C.addressof(n))
    assert nh.STATUS[st] == "invalid"
    assert b"67,108,864" in L.nrt_last_error(None)
