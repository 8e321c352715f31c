"""ctypes binding of the C ABI in include/eio.h (libeio_hip.so, hand-written HIP for gfx950).

This is the Python view of the same boundary the C++ operator layer uses; the
torch FWI layer and the C-ABI tests call through it.  The library must be built
(``__graft_entry__.build()``); a missing library raises instead of falling back.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libeio_hip.so")

EIO_DEVICE_PTRS = 1
VARIANT_AUTO = -1  # EIO_VARIANT_AUTO: two-step tile width chosen per call
GROUP_AUTO = -1    # EIO_GROUP_AUTO: shot groups of a 192 MiB working set
GRID_OF_COMP = (1, 2, 0, 0, 3)


class EioError(RuntimeError):
    pass


class Grid(ctypes.Structure):
    _fields_ = [("nz", ctypes.c_int), ("nx", ctypes.c_int), ("nts", ctypes.c_int),
                ("sub_fd", ctypes.c_int), ("minPad", ctypes.c_int), ("dz", ctypes.c_double),
                ("dx", ctypes.c_double), ("dts", ctypes.c_double), ("alphaCos", ctypes.c_double),
                ("surface", ctypes.c_int)]


class Shot(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int),
                ("src_pos", ctypes.c_void_p * 4), ("n_src", ctypes.c_int * 4),
                ("rec_pos", ctypes.c_void_p * 4), ("n_rec", ctypes.c_int * 4),
                ("src_tr", ctypes.c_void_p * 5), ("rec_tr", ctypes.c_void_p * 5)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EioError("libeio_hip.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.eio_last_error.restype = ctypes.c_char_p
        vp = ctypes.c_void_p
        for name, args in [
            ("eio_ctx_create", [ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]),
            ("eio_ctx_destroy", [vp]),
            ("eio_ctx_set_stream", [vp, vp]),
            ("eio_ctx_set_graph", [vp, ctypes.c_int]),
            ("eio_ctx_graph_info", [vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                                    ctypes.POINTER(ctypes.c_longlong)]),
            ("eio_ctx_last_timing", [vp, ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_longlong),
                                     ctypes.POINTER(ctypes.c_double)]),
            ("eio_grid_check", [ctypes.POINTER(Grid), ctypes.c_int]),
            ("eio_model_add", [vp, ctypes.POINTER(Grid), ctypes.POINTER(vp), ctypes.c_int,
                               ctypes.POINTER(ctypes.c_int)]),
            ("eio_model_set_media", [vp, ctypes.c_int, ctypes.POINTER(vp), ctypes.c_int]),
            ("eio_model_remove", [vp, ctypes.c_int]),
            ("eio_model_set_damping", [vp, ctypes.c_int, vp]),
            ("eio_nl_forward", [vp, ctypes.c_int, ctypes.POINTER(Shot), ctypes.c_int]),
            ("eio_nl_forward_noalloc", [vp, ctypes.c_int, ctypes.POINTER(Shot), ctypes.c_int]),
            ("eio_nl_adjoint", [vp, ctypes.c_int, ctypes.POINTER(Shot), ctypes.c_int]),
            ("eio_born_forward", [vp, ctypes.c_int, ctypes.POINTER(Shot), ctypes.POINTER(vp),
                                  ctypes.c_int]),
            ("eio_born_adjoint", [vp, ctypes.c_int, ctypes.POINTER(Shot), ctypes.POINTER(vp),
                                  ctypes.c_int]),
            ("eio_device_alloc", [vp, ctypes.c_ulonglong, ctypes.POINTER(vp)]),
            ("eio_device_free", [vp, vp]),
            ("eio_memcpy", [vp, vp, vp, ctypes.c_ulonglong, ctypes.c_int]),
            ("eio_synchronize", [vp]),
            ("eio_copy_probe", [vp, vp, vp, ctypes.c_ulonglong, ctypes.c_int, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_double)]),
            ("eio_ctx_set_variant", [vp, ctypes.c_int]),
            ("eio_ctx_variant_info", [ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
            ("eio_ctx_num_variants", []),
            ("eio_fwi_forward", [vp, ctypes.c_int, ctypes.POINTER(Shot), ctypes.c_int, ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_int)]),
            ("eio_fwi_gradient", [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Shot),
                                  ctypes.POINTER(vp), ctypes.c_int]),
            ("eio_fwi_info", [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                              ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_ulonglong)]),
            ("eio_fwi_release", [vp, ctypes.c_int]),
            ("eio_ctx_num_time_block_variants", []),
            ("eio_ctx_time_block_info", [ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
            ("eio_ctx_set_time_block", [vp, ctypes.c_int, ctypes.c_int]),
            ("eio_ctx_num_time_block3_variants", []),
            ("eio_ctx_time_block3_info", [ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
            ("eio_ctx_time_block_layout", [ctypes.c_int]),
            ("eio_ctx_num_adjoint_time_block_variants", []),
            ("eio_ctx_adjoint_time_block_info", [ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                                 ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
            ("eio_ctx_set_adjoint_time_block", [vp, ctypes.c_int, ctypes.c_int]),
            ("eio_ctx_set_step_limit", [vp, ctypes.c_int]),
            ("eio_ctx_set_shot_group_bytes", [vp, ctypes.c_longlong]),
            ("eio_ctx_set_shot_group_streams", [vp, ctypes.c_int]),
            ("eio_ctx_last_shot_groups", [vp, ctypes.POINTER(ctypes.c_int)]),
            ("eio_host_alloc", [ctypes.c_ulonglong, ctypes.POINTER(vp)]),
            ("eio_host_free", [vp]),
            ("eio_ctx_set_persistent", [vp, ctypes.c_int, ctypes.c_int]),
            ("eio_nl_forward_wavefield", [vp, ctypes.c_int, ctypes.POINTER(Shot), ctypes.c_int, ctypes.c_int, vp]),
            ("eio_nl_adjoint_wavefield", [vp, ctypes.c_int, ctypes.POINTER(Shot), ctypes.c_int, ctypes.c_int, vp]),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        _lib = L
    return _lib


EXPORTED = ("eio_last_error", "eio_ctx_create", "eio_ctx_destroy", "eio_ctx_set_stream",
            "eio_ctx_set_graph", "eio_ctx_graph_info", "eio_ctx_last_timing", "eio_grid_check", "eio_model_add", "eio_model_set_media",
            "eio_model_remove", "eio_model_set_damping", "eio_nl_forward", "eio_nl_forward_noalloc", "eio_nl_adjoint", "eio_born_forward",
            "eio_born_adjoint", "eio_device_alloc", "eio_device_free", "eio_memcpy",
            "eio_synchronize", "eio_copy_probe", "eio_ctx_set_variant", "eio_ctx_variant_info",
            "eio_ctx_num_variants", "eio_fwi_forward", "eio_fwi_gradient", "eio_fwi_info",
            "eio_fwi_release", "eio_ctx_num_time_block_variants", "eio_ctx_time_block_info",
            "eio_ctx_set_time_block", "eio_ctx_num_time_block3_variants", "eio_ctx_time_block3_info",
            "eio_ctx_time_block_layout", "eio_ctx_set_step_limit", "eio_ctx_set_persistent", "eio_nl_forward_wavefield",
            "eio_nl_adjoint_wavefield", "eio_model_set_abc", "eio_nl_forward_snapshots",
            "eio_nl_step_states", "eio_ctx_release_workspace", "eio_ctx_num_adjoint_time_block_variants",
            "eio_ctx_adjoint_time_block_info", "eio_ctx_set_adjoint_time_block",
            "eio_ctx_set_shot_group_bytes", "eio_ctx_set_shot_group_streams", "eio_ctx_last_shot_groups",
            "eio_host_alloc", "eio_host_free")


def check(rc):
    if rc != 0:
        raise EioError("eio error %d: %s" % (rc, lib().eio_last_error().decode()))


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


def grid_check(nz, nx, nts, sub_fd, dz, dx, dts, minPad, precision, alphaCos=0.99, surface=0):
    """eio_grid_check: the grid checks of eio_model_add on the host (raises EioError)."""
    g = Grid(int(nz), int(nx), int(nts), int(sub_fd), int(minPad), float(dz), float(dx),
             float(dts), float(alphaCos), int(surface))
    rc = lib().eio_grid_check(ctypes.byref(g), int(precision))
    if rc != 0:
        raise EioError("eio error %d: %s" % (rc, lib().eio_last_error().decode()))
    return rc


class Context:
    """Owns an eio_ctx (one HIP device, fixed real type)."""

    def __init__(self, device=0, dtype=np.float32):
        self.dtype = np.dtype(dtype)
        self.prec = self.dtype.itemsize
        h = ctypes.c_void_p()
        check(lib().eio_ctx_create(device, self.prec, ctypes.byref(h)))
        self.h = h
        self._keep = []

    def close(self):
        if self.h:
            check(lib().eio_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def release_workspace(self):
        """Free the context's cached device workspace (it is kept across calls otherwise)."""
        check(lib().eio_ctx_release_workspace(self.h))

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        check(lib().eio_device_alloc(self.h, max(int(nbytes), 16), ctypes.byref(p)))
        return p.value

    def free(self, ptr):
        check(lib().eio_device_free(self.h, ctypes.c_void_p(ptr)))

    def h2d(self, dptr, arr):
        arr = np.ascontiguousarray(arr)
        check(lib().eio_memcpy(self.h, ctypes.c_void_p(dptr), ctypes.c_void_p(arr.ctypes.data),
                               arr.nbytes, 1))

    def d2h(self, arr, dptr):
        check(lib().eio_memcpy(self.h, ctypes.c_void_p(arr.ctypes.data), ctypes.c_void_p(dptr),
                               arr.nbytes, 2))

    def copy_probe(self, nbytes, reps=10, kind=1):
        """Mean time [ms] of a device copy of nbytes (two buffers allocated here): kind 1 one
        16-byte vector per thread with nontemporal stores, kind 0 a grid-stride copy, kind 2 one
        4-byte word per thread."""
        a, b = self.alloc(nbytes), self.alloc(nbytes)
        ms = ctypes.c_double()
        try:
            check(lib().eio_copy_probe(self.h, ctypes.c_void_p(b), ctypes.c_void_p(a), int(nbytes),
                                       int(reps), int(kind), ctypes.byref(ms)))
        finally:
            self.free(a)
            self.free(b)
        return ms.value

    def set_stream(self, stream_handle):
        check(lib().eio_ctx_set_stream(self.h, ctypes.c_void_p(stream_handle)))

    def set_variant(self, v):
        check(lib().eio_ctx_set_variant(self.h, int(v)))

    @staticmethod
    def variants():
        """[(R, WX, occ)] of the FP32 step-kernel variants."""
        out = []
        for v in range(lib().eio_ctx_num_variants()):
            r, wx, oc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            check(lib().eio_ctx_variant_info(v, ctypes.byref(r), ctypes.byref(wx), ctypes.byref(oc)))
            out.append((r.value, wx.value, oc.value))
        return out

    def set_time_block(self, steps, variant=0):
        """Fine steps per launch of the FP32 NL forward (1, 2 or 3) and the variant of the
        two- or three-step kernel (ignored for 1: the single-step kernel keeps the context's
        own variant); VARIANT_AUTO chooses the tile width per call from the batch's size."""
        check(lib().eio_ctx_set_time_block(self.h, int(steps), int(variant)))

    def set_adjoint_time_block(self, steps, variant=0):
        """Reverse fine steps per launch of the FP32 adjoint passes (1 or 2) and the two-step
        adjoint variant; VARIANT_AUTO chooses the tile width per call from the batch's size."""
        check(lib().eio_ctx_set_adjoint_time_block(self.h, int(steps), int(variant)))

    @staticmethod
    def adjoint_time_block_variants():
        """[(R2, WX, occ)] of the two-step adjoint kernel variants."""
        out = []
        for v in range(lib().eio_ctx_num_adjoint_time_block_variants()):
            r, wx, oc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            check(lib().eio_ctx_adjoint_time_block_info(v, ctypes.byref(r), ctypes.byref(wx), ctypes.byref(oc)))
            out.append((r.value, wx.value, oc.value))
        return out

    def set_persistent(self, mode, spin_limit=0):
        """Two-step NL forward in one persistent launch: 0 off, 1 on, 2 when every tile gets its own
        co-resident workgroup; spin_limit > 0 bounds each neighbour wait (polls)."""
        check(lib().eio_ctx_set_persistent(self.h, int(mode), int(spin_limit)))

    def set_step_limit(self, max_steps):
        """Stop NL loops after max_steps fine steps (0: no limit); for tuning runs."""
        check(lib().eio_ctx_set_step_limit(self.h, int(max_steps)))

    def set_shot_group_bytes(self, nbytes):
        """Working-set budget of the FP32 NL forward's shot groups: GROUP_AUTO (192 MiB, the
        default), 0 (one group) or a byte count; each group runs all its steps before the next."""
        check(lib().eio_ctx_set_shot_group_bytes(self.h, int(nbytes)))

    def set_shot_group_streams(self, n):
        """Streams the shot groups of a call run on (2, the default, or 1)."""
        check(lib().eio_ctx_set_shot_group_streams(self.h, int(n)))

    def last_shot_groups(self):
        """Number of shot groups of the last NL forward call."""
        n = ctypes.c_int()
        check(lib().eio_ctx_last_shot_groups(self.h, ctypes.byref(n)))
        return n.value

    @staticmethod
    def time_block_variants():
        """[(R2, WX, occ)] of the two-step NL forward kernel variants."""
        out = []
        for v in range(lib().eio_ctx_num_time_block_variants()):
            r, wx, oc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            check(lib().eio_ctx_time_block_info(v, ctypes.byref(r), ctypes.byref(wx), ctypes.byref(oc)))
            out.append((r.value, wx.value, oc.value))
        return out

    @staticmethod
    def time_block3_variants():
        """[(R3, WX, occ)] of the three-step NL forward kernel variants (owned tile 48 x WX*R3)."""
        out = []
        for v in range(lib().eio_ctx_num_time_block3_variants()):
            r, wx, oc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            check(lib().eio_ctx_time_block3_info(v, ctypes.byref(r), ctypes.byref(wx), ctypes.byref(oc)))
            out.append((r.value, wx.value, oc.value))
        return out

    @staticmethod
    def time_block_layout(variant):
        """Lane layout of two-step variant `variant`: 0 row lanes with the whole stage, 1 row
        lanes with the streamed stage, 2 four-row quad lanes (16-byte accesses)."""
        return lib().eio_ctx_time_block_layout(int(variant))

    def set_graph(self, on):
        check(lib().eio_ctx_set_graph(self.h, int(bool(on))))

    def graph_info(self):
        """(graphs built, graph launches, kernel nodes updated before a replay)."""
        b, l, u = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        check(lib().eio_ctx_graph_info(self.h, ctypes.byref(b), ctypes.byref(l), ctypes.byref(u)))
        return b.value, l.value, u.value

    def last_timing(self):
        t, n, s = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double()
        check(lib().eio_ctx_last_timing(self.h, ctypes.byref(t), ctypes.byref(n), ctypes.byref(s)))
        return t.value, n.value, s.value

    def model_add(self, nz, nx, nts, sub_fd, dz, dx, dts, minPad, media, alphaCos=0.99,
                  device_ptrs=None, surface=0):
        """media: array (5, nx, nz) on the host, or 5 device addresses (device_ptrs);
        surface=1: free surface at row 2*FAT."""
        g = Grid(int(nz), int(nx), int(nts), int(sub_fd), int(minPad), float(dz), float(dx),
                 float(dts), float(alphaCos), int(surface))
        mid = ctypes.c_int()
        if device_ptrs is not None:
            arr = (ctypes.c_void_p * 5)(*device_ptrs)
            check(lib().eio_model_add(self.h, ctypes.byref(g), arr, EIO_DEVICE_PTRS, ctypes.byref(mid)))
        else:
            med = np.ascontiguousarray(media, dtype=self.dtype)
            arr = (ctypes.c_void_p * 5)(*[med[k].ctypes.data for k in range(5)])
            check(lib().eio_model_add(self.h, ctypes.byref(g), arr, 0, ctypes.byref(mid)))
        return mid.value

    def model_set_media(self, mid, media=None, device_ptrs=None):
        if device_ptrs is not None:
            arr = (ctypes.c_void_p * 5)(*device_ptrs)
            check(lib().eio_model_set_media(self.h, mid, arr, EIO_DEVICE_PTRS))
        else:
            med = np.ascontiguousarray(media, dtype=self.dtype)
            arr = (ctypes.c_void_p * 5)(*[med[k].ctypes.data for k in range(5)])
            check(lib().eio_model_set_media(self.h, mid, arr, 0))

    def set_damping(self, mid, table):
        """Replace the model's sponge table (minPad factors in (0, 1])."""
        t = np.ascontiguousarray(table, self.dtype)
        check(lib().eio_model_set_damping(self.h, int(mid), t.ctypes.data_as(ctypes.c_void_p)))

    def model_remove(self, mid):
        check(lib().eio_model_remove(self.h, mid))

    # -------------------------------------------------------------- shots (host arrays)
    def make_shot(self, mid, spos, rpos, src_tr, rec_tr):
        """spos/rpos: 4 int32 arrays; src_tr/rec_tr: 5 arrays (host numpy) or 5 device addresses."""
        s = Shot()
        s.model = mid
        keep = []
        for g in range(4):
            a = np.ascontiguousarray(spos[g], dtype=np.int32)
            b = np.ascontiguousarray(rpos[g], dtype=np.int32)
            keep += [a, b]
            s.src_pos[g] = a.ctypes.data if a.size else None
            s.n_src[g] = a.size
            s.rec_pos[g] = b.ctypes.data if b.size else None
            s.n_rec[g] = b.size
        for k in range(5):
            for arrname, tr in (("src_tr", src_tr), ("rec_tr", rec_tr)):
                x = tr[k]
                if isinstance(x, np.ndarray):
                    if not x.flags.c_contiguous or x.dtype != self.dtype:
                        raise EioError("trace arrays must be C-contiguous %s" % self.dtype)
                    getattr(s, arrname)[k] = x.ctypes.data if x.size else None
                    keep.append(x)
                else:
                    getattr(s, arrname)[k] = int(x) if x else None
        self._keep = keep
        return s, keep

    def _shots(self, shots):
        arr = (Shot * len(shots))(*shots)
        return arr

    def nl_forward(self, shots, device_ptrs=False):
        arr = self._shots(shots)
        check(lib().eio_nl_forward(self.h, len(shots), arr, EIO_DEVICE_PTRS if device_ptrs else 0))

    def nl_forward_noalloc(self, shots):
        """eio_nl_forward without allocation (device traces; workspace from the library's
        64 MiB per-GPU scratch array)."""
        arr = self._shots(shots)
        check(lib().eio_nl_forward_noalloc(self.h, len(shots), arr, EIO_DEVICE_PTRS))

    def nl_adjoint(self, shots, device_ptrs=False):
        arr = self._shots(shots)
        check(lib().eio_nl_adjoint(self.h, len(shots), arr, EIO_DEVICE_PTRS if device_ptrs else 0))

    def nl_forward_wavefield(self, shots, shot, wavefield, device_ptrs=False):
        """NL forward saving shot `shot`'s wavefield into `wavefield` ((nts, 5, nx, nz) host array
        of the context dtype, or a device address)."""
        arr = self._shots(shots)
        w = ctypes.c_void_p(int(wavefield)) if device_ptrs else self._wfld_ptr(wavefield)
        check(lib().eio_nl_forward_wavefield(self.h, len(shots), arr, EIO_DEVICE_PTRS if device_ptrs else 0,
                                             int(shot), w))

    def nl_adjoint_wavefield(self, shots, shot, wavefield, device_ptrs=False):
        arr = self._shots(shots)
        w = ctypes.c_void_p(int(wavefield)) if device_ptrs else self._wfld_ptr(wavefield)
        check(lib().eio_nl_adjoint_wavefield(self.h, len(shots), arr, EIO_DEVICE_PTRS if device_ptrs else 0,
                                             int(shot), w))

    def set_abc(self, mid, on, weights=None, scale=None):
        """Attach (on = 1: difference inputs, 2: level inputs; + 10 x depth to write only the
        `depth` frame rows next to the update region) or detach (0) a learned boundary:
        weights FP32, the four edges' three layers in torch Conv2d layout (49600 values),
        scale[10] (new level, then differences or previous level)."""
        if not on:
            check(lib().eio_model_set_abc(self.h, int(mid), 0, None, None))
            return
        w = np.ascontiguousarray(weights, dtype=np.float32).reshape(-1)
        if w.size != 4 * (16 * 10 + 16 * 16 + 5 * 16) * 25:
            raise EioError("learned-boundary weights must hold 49600 values")
        sc = np.ascontiguousarray(scale, dtype=np.float32).reshape(-1)
        if sc.size != 10:
            raise EioError("learned-boundary scale must hold 10 values")
        check(lib().eio_model_set_abc(self.h, int(mid), int(on), w.ctypes.data_as(ctypes.c_void_p),
                                      sc.ctypes.data_as(ctypes.c_void_p)))

    def nl_forward_snapshots(self, shots, shot, steps, out_ptr):
        """NL forward copying shot `shot`'s (older, newer) levels after the given fine steps to
        the device buffer out_ptr ([len(steps)][2][5][nx][nz]); traces are device pointers."""
        arr = self._shots(shots)
        st = np.ascontiguousarray(steps, dtype=np.int32)
        check(lib().eio_nl_forward_snapshots(self.h, len(shots), arr, EIO_DEVICE_PTRS, int(shot), int(st.size),
                                             st.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(int(out_ptr))))

    def step_states(self, mid, nstate, states_ptr):
        """One fine step of nstate device states [nstate][2][5][nx][nz] of model mid, in place."""
        check(lib().eio_nl_step_states(self.h, int(mid), int(nstate), ctypes.c_void_p(int(states_ptr)),
                                       EIO_DEVICE_PTRS))

    def _wfld_ptr(self, a):
        if not isinstance(a, np.ndarray) or not a.flags.c_contiguous or a.dtype != self.dtype:
            raise EioError("wavefield must be a C-contiguous %s array" % self.dtype)
        return ctypes.c_void_p(a.ctypes.data)

    def born_forward(self, shots, dm5, device_ptrs=False):
        arr = self._shots(shots)
        if device_ptrs:
            dp = (ctypes.c_void_p * 5)(*dm5)
        else:
            dm5 = np.ascontiguousarray(dm5, dtype=self.dtype)
            dp = (ctypes.c_void_p * 5)(*[dm5[k].ctypes.data for k in range(5)])
        check(lib().eio_born_forward(self.h, len(shots), arr, dp,
                                     EIO_DEVICE_PTRS if device_ptrs else 0))

    def born_adjoint(self, shots, dm5_out, device_ptrs=False):
        """dm5_out (host (5,nx,nz) array, accumulated in place) or 5 device addresses."""
        arr = self._shots(shots)
        if device_ptrs:
            dp = (ctypes.c_void_p * 5)(*dm5_out)
        else:
            if not dm5_out.flags.c_contiguous or dm5_out.dtype != self.dtype:
                raise EioError("dm must be C-contiguous %s" % self.dtype)
            dp = (ctypes.c_void_p * 5)(*[dm5_out[k].ctypes.data for k in range(5)])
        check(lib().eio_born_adjoint(self.h, len(shots), arr, dp,
                                     EIO_DEVICE_PTRS if device_ptrs else 0))

    # -------------------------------------------------------------- checkpointed FWI gradient
    def fwi_forward(self, shots, segment=0, device_ptrs=False):
        """NL forward with checkpoints; returns the store handle."""
        arr = self._shots(shots)
        h = ctypes.c_int()
        check(lib().eio_fwi_forward(self.h, len(shots), arr, EIO_DEVICE_PTRS if device_ptrs else 0,
                                    int(segment), ctypes.byref(h)))
        return h.value

    def fwi_gradient(self, handle, shots, dm5_out, device_ptrs=False):
        arr = self._shots(shots)
        if device_ptrs:
            dp = (ctypes.c_void_p * 5)(*dm5_out)
        else:
            if not dm5_out.flags.c_contiguous or dm5_out.dtype != self.dtype:
                raise EioError("dm must be C-contiguous %s" % self.dtype)
            dp = (ctypes.c_void_p * 5)(*[dm5_out[k].ctypes.data for k in range(5)])
        check(lib().eio_fwi_gradient(self.h, int(handle), len(shots), arr, dp,
                                     EIO_DEVICE_PTRS if device_ptrs else 0))

    def fwi_info(self, handle):
        """(segment length, number of checkpoints, device bytes held)."""
        L, n, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_ulonglong()
        check(lib().eio_fwi_info(self.h, int(handle), ctypes.byref(L), ctypes.byref(n), ctypes.byref(b)))
        return L.value, n.value, b.value

    def fwi_release(self, handle):
        check(lib().eio_fwi_release(self.h, int(handle)))
