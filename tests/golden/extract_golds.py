#!/usr/bin/env python3
"""Extract golden vectors from the reference's unit-test sources into JSON fixtures.

This reads the reference's C++ test files as *text* and copies the numeric
literals of the golden arrays (data only) into ``tests/golden/*.json``.  It is
run in the build container where the reference tree is mounted; the JSON it
writes is committed, so nothing at test time needs the reference.

Sources (paths relative to the reference root):
  unit_tests/kernels/UnitTestScalarDiffElem.C:14-27       heat-conduction 8x8 lhs
  unit_tests/kernels/UnitTestContinuityAdvElem.C:14-57    continuity rhs/lhs (default, reduced sens.)
  unit_tests/kernels/UnitTestMomentumAdvDiffElem.C:14-270 momentum rhs[24], lhs[24][24]
  unit_tests/kernels/UnitTestMomentumMassElem.C:14-150     momentum time derivative (consistent, lumped)
  unit_tests/kernels/UnitTestContinuityMassElem.C:14-96    density time derivative (rhs -12.5 everywhere)
  unit_tests/UnitTestKokkosMEGold.h:13-79                 hex8 SCV grad_op golds (perturbed hex)
  unit_tests/UnitTestUtils.C:360-371                      perturbed hex8 coordinates
  unit_tests/UnitTestTpetra.C:86-114,146-158              Tpetra 1x1x2 element matrix, assembled 12x12, row lengths
  unit_tests/kernels/UnitTestKernelUtils.C:238-271        TrigFieldFunction constants
  unit_tests/UnitTestUtils.C:41-53                        LCG parameters
"""
import json
import os
import re
import sys

REF = os.environ.get("NALU_REFERENCE_ROOT", "/root/reference")
OUT = os.path.dirname(os.path.abspath(__file__))

NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?(?:\s*\*\s*\d+(?:\.\d*)?)?")


def _read(rel):
    with open(os.path.join(REF, rel), "r") as f:
        return f.read()


def _block_after(text, anchor, start=0):
    """Return the text between the '{' that follows `anchor` and its matching '};'."""
    i = text.index(anchor, start)
    j = text.index("=", i)
    k = text.index("{", j)
    depth = 0
    for m in range(k, len(text)):
        c = text[m]
        if c == "{":
            depth += 1
        elif c == "}":
            depth -= 1
            if depth == 0:
                return text[k:m + 1], m
    raise ValueError("unterminated block for %s" % anchor)


def _numbers(block):
    vals = []
    for tok in NUM.findall(block):
        if "*" in tok:  # e.g. "2.0*2" in the Tpetra assembled-matrix gold
            a, b = tok.split("*")
            vals.append(float(a) * float(b))
        else:
            vals.append(float(tok))
    return vals


def _array(text, anchor, start=0, shape=None):
    block, _ = _block_after(text, anchor, start)
    vals = _numbers(block)
    if shape is not None:
        n = 1
        for s in shape:
            n *= s
        if len(vals) != n:
            raise ValueError("%s: expected %d values, got %d" % (anchor, n, len(vals)))
    return vals


def main():
    golds = {}

    t = _read("unit_tests/kernels/UnitTestScalarDiffElem.C")
    golds["scalar_diff"] = {
        "source": "unit_tests/kernels/UnitTestScalarDiffElem.C:14-27,35-65",
        "mesh": "generated:1x1x1 unit cube", "k": 1.0,
        "T": "k/4*(cos(2 aT pi x)+cos(2 aT pi y)+cos(2 aT pi z)), aT=1",
        "rhs": [0.0] * 8,
        "lhs": _array(t, "double lhs[8][8]", shape=(8, 8)),
        "tol": 1e-15,
    }

    t = _read("unit_tests/kernels/UnitTestContinuityAdvElem.C")
    p0 = t.index("namespace advection_default")
    p1 = t.index("namespace advection_reduced_sensitivities")
    golds["continuity_default"] = {
        "source": "unit_tests/kernels/UnitTestContinuityAdvElem.C:16-35,61-100",
        "rhs": _array(t, "double rhs[8]", p0, (8,)),
        "lhs": _array(t, "double lhs[8][8]", p0, (8, 8)),
        "tol": 1e-15,
    }
    golds["continuity_reduced_sens"] = {
        "source": "unit_tests/kernels/UnitTestContinuityAdvElem.C:37-57,105-145",
        "rhs": _array(t, "double rhs[8]", p1, (8,)),
        "lhs": _array(t, "double lhs[8][8]", p1, (8, 8)),
        "tol": 1e-15,
    }

    t = _read("unit_tests/kernels/UnitTestMomentumAdvDiffElem.C")
    golds["momentum_advdiff"] = {
        "source": "unit_tests/kernels/UnitTestMomentumAdvDiffElem.C:17-270,274-306",
        "mu": 0.1, "rho": 1.0, "includeDivU": 0.0,
        "rhs": _array(t, "double rhs[24]", 0, (24,)),
        "lhs": _array(t, "double lhs[24][24]", 0, (24, 24)),
        "tol": 1e-15,
    }

    t = _read("unit_tests/kernels/UnitTestMomentumMassElem.C")
    p0 = t.index("namespace momentum_time_derivative {")
    p1 = t.index("namespace momentum_time_derivative_lumped")
    golds["momentum_mass"] = {
        "source": "unit_tests/kernels/UnitTestMomentumMassElem.C:16-136,140-181",
        "dt": 0.1, "gamma": [1.0, -1.0, 0.0],
        "fields": "rho Np1 = 1, rho N = 0 (unset state), u Np1 = trig velocity, u N = 0, Gjp = trig dpdx",
        "rhs": _array(t, "double rhs[24]", p0, (24,)),
        "lhs": _array(t, "double lhs[24][24]", p0, (24, 24)),
        "tol": 1e-15,
    }
    golds["momentum_mass_lumped"] = {
        "source": "unit_tests/kernels/UnitTestMomentumMassElem.C:138-148,186-225",
        "dt": 0.1, "gamma": [1.0, -1.0, 0.0],
        "rhs": _array(t, "double rhs[24]", p1, (24,)),
        "lhs_diag": 1.25,
        "tol": 1e-15,
    }
    golds["continuity_mass"] = {
        "source": "unit_tests/kernels/UnitTestContinuityMassElem.C:14-54,56-96",
        "dt": 0.1, "gamma": [1.0, -1.0, 0.0],
        "rhs_all": -12.5, "lhs_all": 0.0, "lumped_too": True,
        "tol": 1e-15,
    }

    t = _read("unit_tests/UnitTestKokkosMEGold.h")
    golds["hex8_scv_grad_op"] = {
        "source": "unit_tests/UnitTestKokkosMEGold.h:13-41",
        "values": _array(t, "hex8_scv_grad_op[192]", 0, (192,)),
        "tol": 1e-10,
    }
    golds["hex8_scv_shifted_grad_op"] = {
        "source": "unit_tests/UnitTestKokkosMEGold.h:43-79",
        "values": _array(t, "hex8_scv_shifted_grad_op[192]", 0, (192,)),
        "tol": 1e-10,
    }

    t = _read("unit_tests/UnitTestUtils.C")
    ph = t.index("create_one_perturbed_element")
    h8 = t.index("case stk::topology::HEXAHEDRON_8", ph)
    golds["perturbed_hex8_coords"] = {
        "source": "unit_tests/UnitTestUtils.C:360-371",
        "values": _array(t, "nodeLocations", h8, (8, 3)),
    }

    t = _read("unit_tests/UnitTestTpetra.C")
    golds["tpetra_1x1x2"] = {
        "source": "unit_tests/UnitTestTpetra.C:86-114,146-158,190-201",
        "elem_vals": _array(t, "elemVals[8][8]", 0, (8, 8)),
        "lhs_vals": _array(t, "lhsVals[12][12]", 0, (12, 12)),
        "row_lengths": [int(v) for v in _array(t, "goldRowLengths", 0, (12,))],
        "tol": 1e-9,
    }

    golds["trig_field_constants"] = {
        "source": "unit_tests/kernels/UnitTestKernelUtils.C:238-271",
        "unot": 1.0, "vnot": 1.0, "pnot": 1.0, "k": 1.0, "a": 0.3, "aT": 1.0, "rhonot": 1.0,
    }
    golds["lcg"] = {
        "source": "unit_tests/UnitTestUtils.C:41-53",
        "a": 1103515245, "c": 12345, "m": 2147483647,
    }

    path = os.path.join(OUT, "reference_golds.json")
    with open(path, "w") as f:
        json.dump(golds, f, indent=1)
    print("wrote", path, "with", len(golds), "entries")


if __name__ == "__main__":
    sys.exit(main())
