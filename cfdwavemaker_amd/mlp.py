"""Python handle on the surrogate MLP of libcfdwm (cfdwm_mlp_* C ABI).

Device buffers are torch tensors on the GPU (the library reads and writes
them through their data pointers); parameters and normalisation travel as
float32 numpy arrays.
"""
import ctypes as C

import numpy as np

from ._lib import check, load

_fp = C.POINTER(C.c_float)


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _fptr(a):
    return a.ctypes.data_as(_fp)


class Mlp:
    def __init__(self, nin, width, nhidden, nout, seed=0):
        self.L = load()
        h = C.c_void_p()
        check(self.L.cfdwm_mlp_create(C.byref(h), int(nin), int(width), int(nhidden), int(nout), int(seed)),
              "cfdwm_mlp_create")
        self.h = h
        self.nin, self.width, self.nhidden, self.nout = nin, width, nhidden, nout

    def close(self):
        if getattr(self, "h", None):
            self.L.cfdwm_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def dims(self):
        d, out = self.nin, []
        for l in range(self.nhidden + 1):
            o = self.width if l < self.nhidden else self.nout
            out.append((d, o))
            d = o
        return out

    def nparams(self):
        return int(self.L.cfdwm_mlp_nparams(self.h))

    def get_params(self):
        p = np.zeros(self.nparams(), np.float32)
        check(self.L.cfdwm_mlp_get_params(self.h, _fptr(p)), "cfdwm_mlp_get_params")
        return p

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.nparams()
        check(self.L.cfdwm_mlp_set_params(self.h, _fptr(p)), "cfdwm_mlp_set_params")

    def set_norm(self, in_mean, in_std, out_mean, out_std):
        a = [_f32(x) for x in (in_mean, in_std, out_mean, out_std)]
        check(self.L.cfdwm_mlp_set_norm(self.h, *[_fptr(x) for x in a]), "cfdwm_mlp_set_norm")

    def get_norm(self):
        n = np.zeros(2 * self.nin + 2 * self.nout, np.float32)
        check(self.L.cfdwm_mlp_get_norm(self.h, _fptr(n)), "cfdwm_mlp_get_norm")
        a, b = self.nin, self.nout
        return n[:a], n[a:2 * a], n[2 * a:2 * a + b], n[2 * a + b:]

    def forward(self, X):
        """X: float32 cuda tensor [n, nin] -> float32 cuda tensor [n, nout]."""
        import torch
        X = X.contiguous()
        Y = torch.empty((X.shape[0], self.nout), dtype=torch.float32, device=X.device)
        check(self.L.cfdwm_mlp_forward_device(self.h, X.shape[0], X.data_ptr(), Y.data_ptr()), "cfdwm_mlp_forward_device")
        check(self.L.cfdwm_sync(), "cfdwm_sync")
        return Y

    def loss_grad(self, X, Y):
        loss = C.c_float()
        g = np.zeros(self.nparams(), np.float32)
        check(self.L.cfdwm_mlp_loss_grad_device(self.h, X.shape[0], X.contiguous().data_ptr(),
                                                Y.contiguous().data_ptr(), C.byref(loss), _fptr(g)),
              "cfdwm_mlp_loss_grad_device")
        return loss.value, g

    def train(self, X, Y, steps, batch=4096, lr=1e-3, seed=0, report=50):
        curve = np.zeros(max(steps, 1), np.float32)
        check(self.L.cfdwm_mlp_train_device(self.h, X.shape[0], X.contiguous().data_ptr(), Y.contiguous().data_ptr(),
                                            int(steps), int(batch), float(lr), int(seed), int(report), _fptr(curve)),
              "cfdwm_mlp_train_device")
        return curve[:steps]

    def save(self, path, **extra):
        mi, si, mo, so = self.get_norm()
        np.savez(path, params=self.get_params(), dims=np.array([self.nin, self.width, self.nhidden, self.nout]),
                 in_mean=mi, in_std=si, out_mean=mo, out_std=so, **extra)

    @classmethod
    def load_file(cls, path):
        z = np.load(path, allow_pickle=False)
        nin, width, nhidden, nout = (int(v) for v in z["dims"])
        m = cls(nin, width, nhidden, nout)
        m.set_params(z["params"])
        m.set_norm(z["in_mean"], z["in_std"], z["out_mean"], z["out_std"])
        return m, z
