"""Multi-session batches over the C ABI (kf_batch_*): the frame block of many
sessions, one launch per kernel type, sessions with different volume sizes."""
import ctypes as C

from . import _lib as K


class SessionBatch:
    """A batch of GFusion sessions (allocated, same depth size / ICP schedule /
    storage / options).  `frame(depth_ptrs, active)` runs one frame for every
    active session; results equal running each session alone."""

    def __init__(self, sessions, stream=None):
        self._members = list(sessions)
        arr = (C.c_void_p * len(self._members))(*[g.session.value for g in self._members])
        h = C.c_void_p()
        K.check(K.lib().kf_batch_create(arr, len(self._members), stream, C.byref(h)))
        self._b = h

    def setOption(self, option, value):
        """KF_OPT_GRAPH / KF_OPT_ICP_PERSISTENT (kf_batch_set_option)."""
        K.check(K.lib().kf_batch_set_option(self._b, int(option), int(value)))

    def getOption(self, option):
        v = C.c_int()
        K.check(K.lib().kf_batch_get_option(self._b, int(option), C.byref(v)))
        return v.value

    @property
    def members(self):
        return list(self._members)

    def set_sessions(self, sessions):
        self._members = list(sessions)
        arr = (C.c_void_p * len(self._members))(*[g.session.value for g in self._members])
        K.check(K.lib().kf_batch_set_sessions(self._b, arr, len(self._members)))

    def frame(self, depth_dev_ptrs, active=None):
        """depth_dev_ptrs: device addresses (int) of 512x424 uint16 mm frames,
        one per member (ignored for inactive members)."""
        n = len(self._members)
        ptrs = (C.c_void_p * n)(*[p if p else None for p in depth_dev_ptrs])
        act = None
        if active is not None:
            act = (C.c_int * n)(*[1 if a else 0 for a in active])
        K.check(K.lib().kf_batch_frame_u16_dev(self._b, ptrs, act))

    def stream(self):
        h = C.c_void_p()
        K.check(K.lib().kf_batch_get_stream(self._b, C.byref(h)))
        return h.value

    def synchronize(self):
        K.check(K.lib().kf_batch_synchronize(self._b))

    def close(self):
        if self._b:
            K.lib().kf_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
