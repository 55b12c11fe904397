"""The autotuner's JSON plan cache: reading, validation, clearing (no GPU:
loading without the device check makes no device call)."""
import json

import pytest

import fft_amd
from fft_amd import _lib as L


def _write(path, entries, fmt="sfft-tune-1"):
    path.write_text(json.dumps({"format": fmt, "arch": "gfx950", "hip_runtime": 1,
                                "entries": entries}))
    return str(path)


def test_load_counts_entries_and_clears(tmp_path):
    p = _write(tmp_path / "t.json", [
        {"n": 1920, "kind": 0, "precision": 0, "variant": 1, "radices": [], "tpw": 0,
         "ept": 0, "split": 0, "mfma": -1, "ms": 0.1, "gbs": 4000.0},
        {"n": 65536, "kind": 0, "precision": 0, "variant": -1, "radices": [], "tpw": 0,
         "ept": 0, "split": 256, "mfma": -1, "ms": 0.2, "gbs": 2000.0},
        {"n": 4096, "kind": 0, "precision": 2, "variant": -1, "radices": [16, 16, 16],
         "tpw": 2, "ept": 16, "split": 0, "mfma": 0, "ms": 0.3, "gbs": 900.0},
    ])
    assert fft_amd.load_tuning(p, check_device=False) == 3
    fft_amd.clear_tuning()


@pytest.mark.parametrize("bad", [
    "{",                                   # truncated
    '{"format": "other", "entries": []}',  # unknown format
    '{"format": "sfft-tune-1", "entries": [{"n": 0, "kind": 0, "precision": 0}]}',
    '{"format": "sfft-tune-1", "entries": [{"n": 64, "kind": 7, "precision": 0}]}',
    '{"format": "sfft-tune-1", "entries": [{"n": 64, "kind": 0, "precision": 0, '
    '"radices": [2,2,2,2,2,2,2,2,2,2,2,2,2,2,2,2]}]}',
])
def test_malformed_files_rejected(tmp_path, bad):
    p = tmp_path / "bad.json"
    p.write_text(bad)
    with pytest.raises(L.SfftError) as e:
        fft_amd.load_tuning(str(p), check_device=False)
    assert e.value.code == L.SFFT_ERR_INVALID_ARG


def test_missing_file_rejected(tmp_path):
    with pytest.raises(L.SfftError):
        fft_amd.load_tuning(str(tmp_path / "none.json"), check_device=False)
