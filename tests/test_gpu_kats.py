"""The product against the hand-worked known answers of tests/test_oracle_kats.py
where it has an entry point for them: the calibration matrix counted on the
device (gsnp_matrix_gen) must hit every fall-back branch of cpp/matrix.cc:129-205
exactly as worked out by hand."""
from __future__ import annotations

import pytest

from test_oracle_kats import check_matrix_kat, matrix_kat_inputs

pytestmark = pytest.mark.gpu


def test_device_matrix_gen_fallbacks_by_hand(built, tmp_path):
    fa, so, snp = matrix_kat_inputs(tmp_path)
    print("[kat] device matrix_gen on the hand-counted reads", flush=True)
    check_matrix_kat(built.matrix_gen(fa, so, dbsnp=snp, read_length=4))
