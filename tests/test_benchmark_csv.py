"""The reference-format results writer (no GPU needed)."""
from fft_amd import benchmark_csv as bc


def test_reference_size_list():
    s = bc.reference_sizes()
    assert s[:8] == [1, 2, 3, 4, 6, 8, 9, 12]
    assert s[-1] == 1 << 24 and (3 << 22) in s and (3 << 23) not in s and (9 << 20) in s
    assert s == sorted(s) and len(s) == len(set(s)) == 25 + 23 + 21


def test_csv_and_js_format(tmp_path):
    rows = [(4, 1000.0, 0.0055451774444795625), (1024, 5.0, 0.0354891)]
    bc.write(rows, "GPU", str(tmp_path / "tag"))
    csv = (tmp_path / "tag.csv").read_text().splitlines()
    assert csv[0] == "size,ops/sec,GPU" and csv[1].startswith("4,1000,")
    js = (tmp_path / "tag.js").read_text()
    assert js.startswith('addResults("GPU", [') and js.endswith("]);") and "{size: 1024," in js
