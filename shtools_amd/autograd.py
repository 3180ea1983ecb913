"""Differentiable transforms on device tensors.

The four transforms (and the batched DH pair) as ``torch.autograd.Function``s
whose backward passes run the library's adjoint kernels, which recompute the
Legendre functions instead of storing them.  The adjoints are themselves
differentiable (their backward is the forward transform), so second
derivatives work too.

    make_grid_dh(cilm, lmax=None, norm=1, sampling=1, csphase=1, lmax_calc=None, extend=0)
    sh_expand_dh(grid, norm=1, sampling=1, csphase=1, lmax_calc=None)
    make_grid_glq(cilm, zero, lmax=None, norm=1, csphase=1, lmax_calc=None, extend=0)
    sh_expand_glq(grid, w, zero, norm=1, csphase=1, lmax_calc=None)
    batch_make_grid_dh(batch, cilms) / batch_sh_expand_dh(batch, grids)   (DHBatch plans)

Arguments are those of the backend functions (MakeGridDH etc.); inputs must be
float64 tensors on a HIP device.
"""
import torch

from . import backend as B


def _contig(t):
    return t if t.is_contiguous() else t.contiguous()


class _MakeGridDH(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cilm, args):
        ctx.args = args
        ctx.cshape = tuple(cilm.shape)
        return B.MakeGridDH(_contig(cilm), *args)

    @staticmethod
    def backward(ctx, g):
        return _MakeGridDHAdj.apply(_contig(g), ctx.cshape, ctx.args), None


class _MakeGridDHAdj(torch.autograd.Function):
    @staticmethod
    def forward(ctx, g, cshape, args):
        ctx.args = args
        return B.MakeGridDH_adjoint(_contig(g), cshape, *args)

    @staticmethod
    def backward(ctx, c):
        return _MakeGridDH.apply(_contig(c), ctx.args), None, None


class _SHExpandDH(torch.autograd.Function):
    @staticmethod
    def forward(ctx, grid, args):
        ctx.args = args
        ctx.gshape = tuple(grid.shape)
        return B.SHExpandDH(_contig(grid), *args)

    @staticmethod
    def backward(ctx, c):
        return _SHExpandDHAdj.apply(_contig(c), ctx.gshape, ctx.args), None


class _SHExpandDHAdj(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c, gshape, args):
        ctx.args = args
        return B.SHExpandDH_adjoint(_contig(c), gshape, *args)

    @staticmethod
    def backward(ctx, g):
        return _SHExpandDH.apply(_contig(g), ctx.args), None, None


class _MakeGridGLQ(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cilm, zero, args):
        ctx.zero, ctx.args = zero, args
        ctx.cshape = tuple(cilm.shape)
        return B.MakeGridGLQ(_contig(cilm), zero, *args)

    @staticmethod
    def backward(ctx, g):
        return _MakeGridGLQAdj.apply(_contig(g), ctx.cshape, ctx.zero, ctx.args), None, None


class _MakeGridGLQAdj(torch.autograd.Function):
    @staticmethod
    def forward(ctx, g, cshape, zero, args):
        ctx.zero, ctx.args = zero, args
        return B.MakeGridGLQ_adjoint(_contig(g), cshape, zero, *args)

    @staticmethod
    def backward(ctx, c):
        return _MakeGridGLQ.apply(_contig(c), ctx.zero, ctx.args), None, None, None


class _SHExpandGLQ(torch.autograd.Function):
    @staticmethod
    def forward(ctx, grid, w, zero, args):
        ctx.w, ctx.zero, ctx.args = w, zero, args
        ctx.gshape = tuple(grid.shape)
        return B.SHExpandGLQ(_contig(grid), w, zero, *args)

    @staticmethod
    def backward(ctx, c):
        return _SHExpandGLQAdj.apply(_contig(c), ctx.gshape, ctx.w, ctx.zero, ctx.args), None, None, None


class _SHExpandGLQAdj(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c, gshape, w, zero, args):
        ctx.w, ctx.zero, ctx.args = w, zero, args
        return B.SHExpandGLQ_adjoint(_contig(c), gshape, w, zero, *args)

    @staticmethod
    def backward(ctx, g):
        return _SHExpandGLQ.apply(_contig(g), ctx.w, ctx.zero, ctx.args), None, None, None, None


def make_grid_dh(cilm, lmax=None, norm=1, sampling=1, csphase=1, lmax_calc=None, extend=0):
    """Differentiable MakeGridDH (src/MakeGridDH.f95)."""
    return _MakeGridDH.apply(cilm, (lmax, norm, sampling, csphase, lmax_calc, extend))


def sh_expand_dh(grid, norm=1, sampling=1, csphase=1, lmax_calc=None):
    """Differentiable SHExpandDH (src/SHExpandDH.f95)."""
    return _SHExpandDH.apply(grid, (norm, sampling, csphase, lmax_calc))


def make_grid_glq(cilm, zero, lmax=None, norm=1, csphase=1, lmax_calc=None, extend=0):
    """Differentiable MakeGridGLQ (src/MakeGridGLQ.f95); zero is not differentiated."""
    return _MakeGridGLQ.apply(cilm, zero, (lmax, norm, csphase, lmax_calc, extend))


def sh_expand_glq(grid, w, zero, norm=1, csphase=1, lmax_calc=None):
    """Differentiable SHExpandGLQ (src/SHExpandGLQ.f95); w and zero are not differentiated."""
    return _SHExpandGLQ.apply(grid, w, zero, (norm, csphase, lmax_calc))


# ------------------------------------------------------------ batched DH

class _BatchMakeGrid(torch.autograd.Function):
    @staticmethod
    def forward(ctx, batch, *cilms):
        ctx.batch = batch
        return tuple(batch.makegrid([_contig(c) for c in cilms]))

    @staticmethod
    def backward(ctx, *gs):
        return (None,) + tuple(_BatchMakeGridAdj.apply(ctx.batch, *[_contig(g) for g in gs]))


class _BatchMakeGridAdj(torch.autograd.Function):
    @staticmethod
    def forward(ctx, batch, *gs):
        ctx.batch = batch
        return tuple(batch.makegrid_adjoint([_contig(g) for g in gs]))

    @staticmethod
    def backward(ctx, *cs):
        return (None,) + tuple(_BatchMakeGrid.apply(ctx.batch, *[_contig(c) for c in cs]))


class _BatchExpand(torch.autograd.Function):
    @staticmethod
    def forward(ctx, batch, *grids):
        ctx.batch = batch
        return tuple(batch.expand([_contig(g) for g in grids]))

    @staticmethod
    def backward(ctx, *cs):
        return (None,) + tuple(_BatchExpandAdj.apply(ctx.batch, *[_contig(c) for c in cs]))


class _BatchExpandAdj(torch.autograd.Function):
    @staticmethod
    def forward(ctx, batch, *cs):
        ctx.batch = batch
        return tuple(batch.expand_adjoint([_contig(c) for c in cs]))

    @staticmethod
    def backward(ctx, *gs):
        return (None,) + tuple(_BatchExpand.apply(ctx.batch, *[_contig(g) for g in gs]))


def batch_make_grid_dh(batch, cilms):
    """Differentiable DHBatch.makegrid: list of coefficient tensors -> list of grids."""
    return list(_BatchMakeGrid.apply(batch, *cilms))


def batch_sh_expand_dh(batch, grids):
    """Differentiable DHBatch.expand: list of grids -> list of coefficient tensors."""
    return list(_BatchExpand.apply(batch, *grids))
