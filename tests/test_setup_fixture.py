"""The kinetic setup reproduces tests/golden/setup_digests.json bit for bit: the velocity layout,
cell numbering (RCB / RCM), LU-SGS blocks, per-direction sweep orders, CSR face entries, boundary-face
table, per-cell scalars and case description of every case in tests/setup_digest.py. The numbering
and sweep orders fix every multi-block iterate, so a change in any of them changes results that the
parity tests would otherwise only see through the GPU and oracle runs sharing the same setup."""

from __future__ import annotations

import json

import pytest

import setup_digest


@pytest.mark.parametrize("name", sorted(setup_digest.CASES))
def test_setup_tables_match_fixture(gsk, meshgen, tmp_path, name):
    with open(setup_digest.FIXTURE) as f:
        expected = json.load(f)[name]
    got = setup_digest.build_digests(gsk, meshgen, str(tmp_path), names={name})[name]
    assert got["sizes"] == expected["sizes"]
    bad = sorted(k for k in expected if got.get(k) != expected[k])
    assert not bad, f"{name}: tables differ from the fixture: {bad}"
