"""Separable coefficient expressions on the host: algebra, torch evaluation
against the same formula written out, and the layout of the device table
(fa_sep_t of include/fealpy_amd.h).  No GPU."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

from fealpy_amd.expr import FaSep, Sep, cell, cos, exp, sin, x, y, z

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PI = np.pi


def test_expression_matches_written_formula():
    g = torch.Generator().manual_seed(5)
    k = torch.rand(7, dtype=torch.float64, generator=g)
    p = torch.rand(7, 5, 3, dtype=torch.float64, generator=g)
    f = 3 * PI ** 2 * cell(k) * sin(PI * x) * sin(PI * y) * sin(PI * z) + 0.5 * cos(2 * PI * x) * z ** 2 \
        - (x - 1) * exp(0.5 * y + 0.25) + 2.0
    ref = (3 * PI ** 2 * k[:, None] * torch.sin(PI * p).prod(-1) + 0.5 * torch.cos(2 * PI * p[..., 0]) * p[..., 2] ** 2
           - (p[..., 0] - 1) * torch.exp(0.5 * p[..., 1] + 0.25) + 2.0)
    np.testing.assert_allclose(f(p).numpy(), ref.numpy(), rtol=1e-14, atol=1e-14)
    assert f.coordtype == 'cartesian'
    # affine factors stay factors: (1 + x)(2 - y) is one term
    h = (1 + x) * (2 - y)
    assert len(h.terms) == 1 and len(h.terms[0][2]) == 2
    q = torch.rand(11, 2, dtype=torch.float64, generator=g)
    np.testing.assert_allclose(h(q).numpy(), ((1 + q[:, 0]) * (2 - q[:, 1])).numpy(), rtol=1e-15)


def test_expression_limits_and_errors():
    with pytest.raises(TypeError):
        sin(x * y)                     # not affine in one coordinate
    with pytest.raises(ValueError):
        cell(torch.ones(3, dtype=torch.float64)) * cell(torch.ones(3, dtype=torch.float64))
    with pytest.raises(ValueError):
        Sep([(1.0, None, ())] * 17)
    f = cell(torch.ones(4, dtype=torch.float64)) * x
    with pytest.raises(ValueError):
        f(torch.rand(3, 2, dtype=torch.float64))   # the cell axis does not match the field


def test_device_table_layout_matches_header():
    """ctypes offsets of FaSep equal the C layout of fa_sep_t: int32 fields,
    then doubles (8-aligned), then pointers, with the header's capacities."""
    src = open(os.path.join(ROOT, "include", "fealpy_amd.h")).read()
    nt = int(re.search(r"#define FA_SEP_MAX_TERMS (\d+)", src).group(1))
    nf = int(re.search(r"#define FA_SEP_MAX_FACTORS (\d+)", src).group(1))
    ints = 2 + nt + 2 * nf
    assert FaSep.coef.offset == 4 * ints and (4 * ints) % 8 == 0
    assert FaSep.a.offset == FaSep.coef.offset + 8 * nt
    assert FaSep.field.offset == FaSep.b.offset + 8 * nf
    assert FaSep.bcs.offset == FaSep.field.offset + 8 * nt
    assert ctypes.sizeof(FaSep) == FaSep.bcs.offset + 8
