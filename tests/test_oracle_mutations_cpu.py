"""Planted mistakes in the oracle must be caught by the closed-form pins.

Each case copies oracle/pf_oracle.c, replaces one expression with a plausible slip (a gate
constant, sigma^2 for 2 sigma^2, the bearing sign, the order of a quaternion product, the
direction of the world-to-body map, a halved angle ...), builds the copy with the oracle's own
flags and runs every check of test_oracle_closed_forms_cpu.py on it: at least one must fail.
CPU only; the GPU library is not involved.
"""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

import oracle as O
from test_oracle_closed_forms_cpu import ALL_CHECKS

SRC = os.path.join(os.path.dirname(O.__file__), "pf_oracle.c")
MLP = os.path.join(os.path.dirname(O.__file__), "mlp_oracle.c")
CFLAGS = ["-O1", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-std=gnu11", "-shared"]

# (name, expression in pf_oracle.c, the slip)
MUTATIONS = [
    ("range gate value -1.5 g^2", "er * er > rng_gate2) ? -0.5 * gate * gate", "er * er > rng_gate2) ? -1.5 * gate * gate"),
    ("bearing gate value -1.5 g^2", "eb * eb > brg_gate2) ? -0.5 * gate * gate", "eb * eb > brg_gate2) ? -1.5 * gate * gate"),
    ("Cartesian gate value -0.5 g^2", "-0.5 * gate * gate * 3.0", "-0.5 * gate * gate"),
    ("range 1/sr^2 for 1/(2 sr^2)", "inv2rr = 1.0 / (2.0 * sr * sr)", "inv2rr = 1.0 / (sr * sr)"),
    ("bearing 1/sb^2 for 1/(2 sb^2)", "inv2bb = 1.0 / (2.0 * sb * sb)", "inv2bb = 1.0 / (sb * sb)"),
    ("bearing sign flipped", "double dot = (-pred[0] / rn) * vx + (-pred[1] / rn) * vy + (-pred[2] / rn) * vz;",
     "double dot = (pred[0] / rn) * vx + (pred[1] / rn) * vy + (pred[2] / rn) * vz;"),
    ("Cartesian gate threshold without the 3", "cart_gate2 = g2 * 3.0 * sr * sr", "cart_gate2 = g2 * sr * sr"),
    ("range gate threshold with a 3", "rng_gate2 = g2 * sr * sr", "rng_gate2 = g2 * 3.0 * sr * sr"),
    ("range gate on |e| instead of e^2", "er * er > rng_gate2", "er > rng_gate2"),
    ("body frame by R instead of R^T", "quat qc = {f->qw[i], -f->qx[i], -f->qy[i], -f->qz[i]};",
     "quat qc = {f->qw[i], f->qx[i], f->qy[i], f->qz[i]};"),
    ("predict q_odo * q", "qmul(qmul(q, qo), dq)", "qmul(qmul(qo, q), dq)"),
    ("predict q * dq * q_odo", "qmul(qmul(q, qo), dq)", "qmul(qmul(q, dq), qo)"),
    ("predict translation not rotated", "qrot(q, body, wd);", "wd[0] = body[0]; wd[1] = body[1]; wd[2] = body[2];"),
    ("reinit dq * q0", "quat q = qmul(q0, dq);", "quat q = qmul(dq, q0);"),
    ("rotation noise by the full angle", "or_det_cos_sinc(0.25 * ang2, &c, &s);", "or_det_cos_sinc(ang2, &c, &s);"),
    ("softmax without the max shift", "f->w[i] = st(f, exp(f->ll[i] - mx) / sum);", "f->w[i] = st(f, exp(f->ll[i]) / sum);"),
    ("attitude log angle halved", "double ang = 2.0 * atan2(hv, d.w);", "double ang = atan2(hv, d.w);"),
    ("attitude residual q * mean^-1", "quat d = qmul(mc, qi);", "quat d = qmul(qi, mc);"),
    ("ESS from sum w", "out[13] += w * w;", "out[13] += w;"),
]


def _build(args):
    i, (name, old, new), tmp = args
    text = open(SRC).read()
    assert text.count(old) == 1, f"{name}: expression not found exactly once"
    d = os.path.join(tmp, f"m{i}")
    os.makedirs(d)
    src = os.path.join(d, "pf_oracle.c")
    with open(src, "w") as fh:
        fh.write(text.replace(old, new))
    shutil.copy(MLP, d)
    out = os.path.join(d, "liboracle_m.so")
    subprocess.run(["gcc", *CFLAGS, "-o", out, src, os.path.join(d, "mlp_oracle.c"), "-lm"], check=True)
    return out


@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("oracle_mutants"))
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        return list(ex.map(_build, [(i, m, tmp) for i, m in enumerate(MUTATIONS)]))


def _caught(L):
    for check in ALL_CHECKS:
        try:
            check(L)
        except AssertionError:
            return check.__name__
    return None


@pytest.mark.parametrize("idx", range(len(MUTATIONS)), ids=[m[0] for m in MUTATIONS])
def test_planted_mistake_is_caught(mutants, idx):
    L = O.load(mutants[idx])
    assert _caught(L) is not None, f"closed-form pins miss: {MUTATIONS[idx][0]}"


def test_unmutated_copy_passes(tmp_path):
    """The same build recipe on the untouched source passes every check (the harness itself is sound)."""
    out = str(tmp_path / "liboracle_plain.so")
    subprocess.run(["gcc", *CFLAGS, "-o", out, SRC, MLP, "-lm"], check=True)
    assert _caught(O.load(out)) is None
