"""The algorithmic flop constants are recounted from source, not retyped."""

from __future__ import annotations

import os
import re

from conftest import ROOT

from kerrray_amd import flops
from kerrray_amd.geodesics import DP_A, DP_E


def _body(src: str, signature: str) -> str:
    start = src.index(signature)
    brace = src.index("{", start)
    depth, i = 0, brace
    while True:
        if src[i] == "{":
            depth += 1
        elif src[i] == "}":
            depth -= 1
            if depth == 0:
                return src[brace + 1 : i]
        i += 1


def _binary_ops(code: str) -> int:
    """Binary + - * / in assignments (unary minus and index/loop arithmetic excluded)."""
    n = 0
    for line in code.splitlines():
        line = line.split("//")[0]
        if "=" not in line or line.strip().startswith(("for", "if", "#")):
            continue
        rhs = line.split("=", 1)[1]
        rhs = re.sub(r"\[[^\]]*\]", "", rhs)  # indices
        rhs = re.sub(r"static_cast<\w+>", "", rhs)
        rhs = re.sub(r"std::\w+\(", "(", rhs)
        tokens = re.findall(r"[A-Za-z_0-9.]+|[-+*/()]", rhs)
        prev = None
        for t in tokens:
            if t in "+-*/" and prev is not None and prev not in "+-*/(":
                n += 1
            prev = t
    return n


def test_rhs_flops_recounted_from_the_oracle_source() -> None:
    src = open(os.path.join(ROOT, "oracle", "kr_oracle.cpp")).read()
    metric = _body(src, "void inverse_metric_terms(")
    rhs = _body(src, "void geodesic_rhs_scalar(")
    assert metric.count("/ ") == 17, "17 divides in the metric"
    assert _binary_ops(metric) == flops.RHS_METRIC
    assert _binary_ops(rhs) == flops.RHS_EQUATIONS


def test_stage_sum_flops_from_the_tableau_pattern() -> None:
    # Rows 1..5 of A build stage inputs y + h * sum(a_ij k_j); row 6 (y_new) skips its zero;
    # the error estimate h * sum(E_i k_i) skips E_1 = 0.
    per_component = 0
    for row in DP_A[1:6]:
        per_component += 2 * len(row) + 1
    nz6 = sum(1 for v in DP_A[6] if v != 0.0)
    per_component += 2 * nz6 + 1
    nz_e = sum(1 for v in DP_E if v != 0.0)
    per_component += 2 * nz_e
    assert 8 * per_component == flops.STAGE_SUMS
