"""The C ABI library loads and exports every entry point declared in
include/dada2_amd.h; its host-only helpers agree with the oracle.  No GPU."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "dada2_amd.h").read_text()
    decl = r"^(?:int|void\*|uint64_t|const char\*)\s+(d2g_[a-z0-9_]+)\("
    return sorted(set(re.findall(decl, text, flags=re.M)))


def test_library_exports_every_declared_symbol(pkg):
    lib = ctypes.CDLL(str(pkg.LIB_PATH))
    names = declared_symbols()
    assert len(names) >= 14
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(pkg.exported_symbols()) == names


def test_cell_count_matches_oracle(pkg, oracle):
    for l1, l2, band in [(250, 250, 16), (240, 260, 16), (260, 240, 16), (1450, 1480, 32),
                         (17, 16, 3), (40, 55, -1), (1, 1, 16), (300, 299, 64)]:
        assert pkg.vectorized_cells(l1, l2, band) == oracle.vectorized_cells(l1, l2, band)


def test_supported_geometries(pkg):
    assert pkg.vectorized_supported(250, 250, 16)
    assert pkg.vectorized_supported(240, 260, 16)
    assert pkg.vectorized_supported(1450, 1480, 16)
    # wider than one wavefront: the one-pair kernel
    assert pkg.vectorized_supported(300, 299, -1)
    assert pkg.vectorized_supported(250, 120, 16)
    assert pkg.vectorized_supported(3500, 3500, -1)
    # beyond the vectorized aligner's length limit (nwalign.rs:1383)
    assert not pkg.vectorized_supported(3501, 3400, 16)


def test_status_strings(pkg):
    lib = pkg._lib()
    for code in (0, -1, -2, -3, -4, -5, -6):
        assert lib.d2g_status_string(code)
