"""On-disk formats of the host layer, no GPU needed: the ROS map_server PGM + YAML pair
(global_loc.rs:32-130) and the landmark text (landmark_map.rs:59-80).

tests/golden/demo_map.yaml is the reference's landmark fixture (config/demo_map.yaml, data
only).  Expected values are computed here independently (numpy for the occupancy rule, a
regular expression for the landmark lines)."""
import re

import numpy as np
import pytest

from cuda_pf_localization_amd import load_landmarks, load_map_pgm_yaml

GOLDEN = __file__.rsplit("/", 1)[0] + "/golden/"


def pgm_bytes(px, comment=True):
    h, w = px.shape
    head = b"P5\n" + (b"# CREATOR: test\n" if comment else b"") + f"{w} {h}\n255\n".encode()
    return head + px.astype(np.uint8).tobytes()


def test_landmark_text_reference_cases():
    # landmark_map.rs:96-112 (parse_map_text, parse_rejects_garbage)
    m = load_landmarks("# comment\n0: [1, 2, 3]\n5: [4.0, 5.0, 6.0]\n")
    assert set(m) == {0, 5} and m[5][2] == 6.0
    with pytest.raises(ValueError, match="line 1"):
        load_landmarks("0 [1, 2]\n")


def test_landmark_text_forms_and_errors():
    m = load_landmarks("3 1.5 -2 0.25  # plain form\n\n  7: [1e1, 2, 3]\n3: [9, 9, 9]\n")
    assert m == {3: (9.0, 9.0, 9.0), 7: (10.0, 2.0, 3.0)}  # a repeated id replaces
    with pytest.raises(ValueError, match="line 2: invalid"):
        load_landmarks("0: [1, 2, 3]\n1: [a, 2, 3]\n")
    with pytest.raises(ValueError, match="non-negative"):
        load_landmarks("-1: [1, 2, 3]\n")


def test_demo_map_fixture():
    text = open(GOLDEN + "demo_map.yaml").read()
    m = load_landmarks(GOLDEN + "demo_map.yaml")
    want = {}
    for line in text.splitlines():
        g = re.match(r"\s*(\d+):\s*\[([^,]+),([^,]+),([^\]]+)\]", line)
        if g:
            want[int(g.group(1))] = tuple(float(g.group(k)) for k in (2, 3, 4))
    assert len(want) == 30 and m == want


def test_pgm_yaml_occupancy_rule_and_row_flip():
    rng = np.random.default_rng(3)
    px = rng.integers(0, 256, (7, 11))
    px[0, 0], px[0, 1] = 89, 90  # (255 - 89)/255 = 0.651 > 0.65 ; (255 - 90)/255 = 0.647
    yaml = "image: m.pgm\nresolution: 0.05 # m/cell\norigin: [-1.5, 2.0, 0.0]\nnegate: 0\nfree_thresh: 0.196\n"
    cells, res, org = load_map_pgm_yaml(pgm_bytes(px), yaml)
    want = ((255 - px) / 255.0 > 0.65).astype(np.uint8)[::-1]
    assert cells.shape == (7, 11) and np.array_equal(cells, want)
    assert res == 0.05 and org == (-1.5, 2.0)
    assert cells[6, 0] == 1 and cells[6, 1] == 0  # top file row is the last map row
    cells2, _, _ = load_map_pgm_yaml(pgm_bytes(px, comment=False), yaml + "occupied_thresh: 0.9\n")
    assert np.array_equal(cells2, ((255 - px) / 255.0 > 0.9).astype(np.uint8)[::-1])


def test_pgm_yaml_errors():
    px = np.zeros((2, 3))
    yaml = "resolution: 0.1\norigin: [0, 0, 0]\n"
    with pytest.raises(ValueError, match="expected P5"):
        load_map_pgm_yaml(b"P2\n3 2\n255\n" + bytes(6), yaml)
    with pytest.raises(ValueError, match="3 x 2 px, body has 4"):
        load_map_pgm_yaml(pgm_bytes(px)[:-2], yaml)  # truncated body
    # dimensions whose product wraps a 64-bit size (2^32 x 2^32) or exceeds int are rejected
    # before any pixel is read; zero dimensions likewise
    for dims in (b"4294967296 4294967296", b"2147483648 1", b"1 2147483648", b"0 5", b"5 0"):
        with pytest.raises(ValueError, match=r"\[1, INT_MAX\]"):
            load_map_pgm_yaml(b"P5\n" + dims + b"\n255\n" + bytes(64), yaml)
    with pytest.raises(ValueError, match="body has 64"):
        load_map_pgm_yaml(b"P5\n2147483647 2147483647\n255\n" + bytes(64), yaml)
    with pytest.raises(ValueError, match="no resolution"):
        load_map_pgm_yaml(pgm_bytes(px), "origin: [0, 0, 0]\n")
    with pytest.raises(ValueError, match="no origin"):
        load_map_pgm_yaml(pgm_bytes(px), "resolution: 0.1\n")
