"""Dense forward-Laplacian arrays on the CPU (oracle; test infrastructure only).

An :class:`FL` holds a value ``x``, its Jacobian ``j`` with respect to ``T`` input
coordinates (leading axis) and its Laplacian ``l`` (the sum of the ``T`` pure second
derivatives).  Arithmetic follows the forward-Laplacian propagation rules that folx 0.2.15
implements for the reference (used at new_sparse_model.py:120,648-656, orbitals.py:150-151,
wave_function.py:178, jastrow.py:399-457):

    (a*b)'' = a''b + ab'' + 2 sum_t a_t b_t,      f(a)'' = f'(a) a'' + f''(a) sum_t a_t^2.

Mixing an FL with a plain tensor treats the tensor as a constant.
"""
from __future__ import annotations

import torch


class FL:
    __slots__ = ("x", "j", "l")

    def __init__(self, x, j, l):
        self.x, self.j, self.l = x, j, l

    # construction --------------------------------------------------------------------
    @staticmethod
    def input(x: torch.Tensor) -> "FL":
        """Independent variables: ``x`` of shape [..., 3] with T = x.numel() tangents."""
        T = x.numel()
        j = torch.eye(T, dtype=x.dtype).reshape(T, *x.shape)
        return FL(x, j, torch.zeros_like(x))

    @property
    def T(self):
        return self.j.shape[0]

    @property
    def shape(self):
        return self.x.shape

    @property
    def dtype(self):
        return self.x.dtype

    # helpers -------------------------------------------------------------------------
    def _jb(self, t):
        # broadcast a per-element tensor against the jacobian (leading tangent axis)
        return t

    def __getitem__(self, idx):
        if not isinstance(idx, tuple):
            idx = (idx,)
        return FL(self.x[idx], self.j[(slice(None),) + idx], self.l[idx])

    def reshape(self, *shape):
        return FL(self.x.reshape(*shape), self.j.reshape(self.T, *shape), self.l.reshape(*shape))

    def unsqueeze(self, dim):
        d = dim if dim >= 0 else self.x.dim() + 1 + dim
        return FL(self.x.unsqueeze(d), self.j.unsqueeze(d + 1), self.l.unsqueeze(d))

    def sum(self, dim):
        d = dim if dim >= 0 else self.x.dim() + dim
        return FL(self.x.sum(d), self.j.sum(d + 1), self.l.sum(d))

    def pad_tangents(self, T_total, offset):
        j = torch.zeros((T_total,) + tuple(self.x.shape), dtype=self.x.dtype)
        j[offset:offset + self.T] = self.j
        return FL(self.x, j, self.l)

    # arithmetic ----------------------------------------------------------------------
    def __add__(self, o):
        if isinstance(o, FL):
            return FL(self.x + o.x, self.j + o.j, self.l + o.l)
        return FL(self.x + o, self.j + torch.zeros_like(self.x + o), self.l + torch.zeros_like(self.x + o))

    __radd__ = __add__

    def __neg__(self):
        return FL(-self.x, -self.j, -self.l)

    def __sub__(self, o):
        return self + (-o)

    def __rsub__(self, o):
        return (-self) + o

    def __mul__(self, o):
        if isinstance(o, FL):
            x = self.x * o.x
            j = self.j * o.x + o.j * self.x
            l = self.l * o.x + self.x * o.l + 2 * (self.j * o.j).sum(0)
            return FL(x, j, l)
        return FL(self.x * o, self.j * o, self.l * o)

    __rmul__ = __mul__

    def __truediv__(self, o):
        if isinstance(o, FL):
            return self * reciprocal(o)
        return self * (1.0 / o)

    def __rtruediv__(self, o):
        return reciprocal(self) * o

    def __matmul__(self, W):
        return FL(self.x @ W, self.j @ W, self.l @ W)


def elementwise(a: FL, f, df, d2f) -> FL:
    fx, d1, d2 = f(a.x), df(a.x), d2f(a.x)
    return FL(fx, a.j * d1, a.l * d1 + d2 * (a.j ** 2).sum(0))


def _is_fl(a):
    return isinstance(a, FL)


def sigmoid_(x):
    return torch.sigmoid(x)


def silu(a):
    if not _is_fl(a):
        return a * torch.sigmoid(a)

    def f(x):
        return x * torch.sigmoid(x)

    def df(x):
        s = torch.sigmoid(x)
        return s * (1 + x * (1 - s))

    def d2f(x):
        s = torch.sigmoid(x)
        return s * (1 - s) * (2 + x * (1 - 2 * s))

    return elementwise(a, f, df, d2f)


def exp(a):
    if not _is_fl(a):
        return torch.exp(a)
    return elementwise(a, torch.exp, torch.exp, torch.exp)


def log1p(a):
    if not _is_fl(a):
        return torch.log1p(a)
    return elementwise(a, torch.log1p, lambda x: 1 / (1 + x), lambda x: -1 / (1 + x) ** 2)


def sqrt(a):
    if not _is_fl(a):
        return torch.sqrt(a)
    return elementwise(a, torch.sqrt, lambda x: 0.5 / torch.sqrt(x), lambda x: -0.25 * x ** -1.5)


def reciprocal(a):
    if not _is_fl(a):
        return 1 / a
    return elementwise(a, lambda x: 1 / x, lambda x: -1 / x ** 2, lambda x: 2 / x ** 3)


def square(a):
    if not _is_fl(a):
        return a * a
    return elementwise(a, lambda x: x * x, lambda x: 2 * x, lambda x: 2 * torch.ones_like(x))


def log_abs(a):
    if not _is_fl(a):
        return torch.log(torch.abs(a))
    return elementwise(a, lambda x: torch.log(torch.abs(x)), lambda x: 1 / x, lambda x: -1 / x ** 2)


def cutoff_poly(d, p=4):
    """cutoff_function (model/utils.py:103-112): 1 + a d^p + b d^(p+1) + c d^(p+2) for d<1,
    else 0."""
    a = -(p + 1) * (p + 2) * 0.5
    b = p * (p + 2)
    c = -p * (p + 1) * 0.5

    def f(x):
        return torch.where(x < 1, 1 + a * x ** p + b * x ** (p + 1) + c * x ** (p + 2), torch.zeros_like(x))

    if not _is_fl(d):
        return f(d)

    def df(x):
        v = a * p * x ** (p - 1) + b * (p + 1) * x ** p + c * (p + 2) * x ** (p + 1)
        return torch.where(x < 1, v, torch.zeros_like(x))

    def d2f(x):
        v = a * p * (p - 1) * x ** (p - 2) + b * (p + 1) * p * x ** (p - 1) + c * (p + 2) * (p + 1) * x ** p
        return torch.where(x < 1, v, torch.zeros_like(x))

    return elementwise(d, f, df, d2f)


def norm_last(a):
    """Euclidean norm over the last axis: sqrt(sum x^2)."""
    if not _is_fl(a):
        return torch.sqrt((a * a).sum(-1))
    return sqrt(square(a).sum(-1))


def concat(items, dim):
    """Concatenate FL and constant tensors along an x-axis ``dim`` (non-negative)."""
    T = next(i.T for i in items if _is_fl(i))
    xs, js, ls = [], [], []
    for it in items:
        if _is_fl(it):
            xs.append(it.x), js.append(it.j), ls.append(it.l)
        else:
            xs.append(it)
            js.append(torch.zeros((T,) + tuple(it.shape), dtype=it.dtype))
            ls.append(torch.zeros_like(it))
    return FL(torch.cat(xs, dim), torch.cat(js, dim + 1), torch.cat(ls, dim))


def linear(a, W, b=None):
    y = a @ W
    if b is not None:
        y = y + b
    return y
