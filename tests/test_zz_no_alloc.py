"""pmx_compute_errors_noalloc (fixed per-device scratch, no allocation)
against pmx_compute_errors through the C ABI: bitwise the same four norms on
two calls in a row, the second on a smaller space; the host-side checks that
run before any kernel."""
import ctypes

import numpy as np
import pytest
import torch

from mymoonmd_amd import _lib
from mymoonmd_amd import parmoon as pm
from mymoonmd_amd._lib import PmxError

pytestmark = pytest.mark.gpu

B3 = (1 / 14 ** 0.5, 2 / 14 ** 0.5, 3 / 14 ** 0.5)


def errors(name, ctx, space, params, u):
    assert u.is_cuda and u.dtype == torch.float64 and u.numel() == space.n_dof
    torch.cuda.synchronize()
    out = (ctypes.c_double * 4)()
    _lib.call(name, ctx.h, space.h, ctypes.byref(params), ctypes.c_void_p(u.data_ptr()), out)
    return list(out)


def test_noalloc_errors_equal_the_allocating_entry_point(ctx):
    cases = [("Default_UnitCube_Hexa", 5, 2, pm.cd_params("exp_sin_3d", disc="supg", eps=1e-3, b=B3,
                                                          hexa_quad_rule=2)),
             ("UnitSquare", 3, 1, pm.cd_params("sine_2d", eps=1.0))]
    rng = np.random.default_rng(9)
    for geo, nref, order, prm in cases:
        print("case", geo, nref, order, flush=True)
        d = pm.TDomain(geo, nref)
        s = pm.TFESpace(d, nref, order, ctx)
        u = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
        pm.interpolate(ctx, s, prm, u)
        u += torch.as_tensor(1e-3 * rng.standard_normal(s.n_dof)).cuda()
        # the quadrature tables are put on the device by the first assembly
        with pytest.raises(PmxError) as e:
            errors("pmx_compute_errors_noalloc", ctx, s, prm, u)
        assert e.value.code == 6
        A = pm.FEMatrix(ctx, s)
        f = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
        pm.assemble_cd(ctx, s, prm, A, f)
        ref = errors("pmx_compute_errors", ctx, s, prm, u)
        got = errors("pmx_compute_errors_noalloc", ctx, s, prm, u)
        assert got == ref, (got, ref)
        assert all(v > 0 for v in ref)
        # a second call on the same space gives the same bits
        assert errors("pmx_compute_errors_noalloc", ctx, s, prm, u) == ref
