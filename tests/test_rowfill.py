"""Host check of the lattice row fill (pt-icm_amd/csrc/rowfill.hpp) that the
bit-set cluster growth uses on 2D rows of up to four words: compiled for the
CPU with clang, it must equal the site-by-site closure (a bit joins when a
neighbour along the row, wrapped at L, is in the set and the bit is in D)
for every row length 1..128 on random rows, and on hand-made rows whose
runs cross word boundaries and the wrap."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HDR_DIR = os.path.join(HERE, "..", "pt-icm_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include <cstdlib>
#define PTX_ROWFILL_FN inline
#include "rowfill.hpp"

static uint64_t st = 0x9E3779B97F4A7C15ull;
static uint64_t rnd() { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return st; }

// site-by-site closure: the brute force the fill must equal
static void closure(uint32_t* v, const uint32_t* d, int L) {
    for (bool ch = true; ch;) {
        ch = false;
        for (int i = 0; i < L; ++i) {
            if (!((v[i >> 5] >> (i & 31)) & 1u)) continue;
            for (int dx = -1; dx <= 1; dx += 2) {
                const int j = (i + dx + L) % L;
                if (((d[j >> 5] >> (j & 31)) & 1u) && !((v[j >> 5] >> (j & 31)) & 1u)) {
                    v[j >> 5] |= 1u << (j & 31);
                    ch = true;
                }
            }
        }
    }
}

static void fill(uint32_t* v, const uint32_t* d, int L) {
    switch ((L + 31) / 32) {
        case 1: ptx::fill_row<1>(v, d, L); break;
        case 2: ptx::fill_row<2>(v, d, L); break;
        case 3: ptx::fill_row<3>(v, d, L); break;
        default: ptx::fill_row<4>(v, d, L); break;
    }
}

static long check(const uint32_t* d, const uint32_t* v0, int L) {
    uint32_t a[4], b[4];
    for (int k = 0; k < 4; ++k) a[k] = b[k] = v0[k];
    fill(a, d, L);
    closure(b, d, L);
    for (int k = 0; k < 4; ++k)
        if (a[k] != b[k]) return 1;
    return 0;
}

static void set(uint32_t* x, int i) { x[i >> 5] |= 1u << (i & 31); }

int main() {
    long bad = 0, tot = 0;
    for (int L = 1; L <= 128; ++L) {
        for (int it = 0; it < 4000; ++it) {
            uint32_t d[4] = {0, 0, 0, 0}, v[4] = {0, 0, 0, 0};
            const unsigned pd = 1 + it % 7, pv = it % 5;  // D density (pd+1)/8, seed rate ~pv/16
            for (int i = 0; i < L; ++i)
                if ((rnd() & 7u) <= pd) {
                    set(d, i);
                    if ((rnd() & 15u) < pv) set(v, i);
                }
            bad += check(d, v, L);
            ++tot;
        }
        // one run through the wrap seeded at either end, and D all ones
        uint32_t d[4] = {0, 0, 0, 0}, v[4] = {0, 0, 0, 0}, w[4] = {0, 0, 0, 0};
        for (int i = 0; i < L; ++i)
            if (i < L / 3 || i >= L - L / 3) set(d, i);
        set(v, 0);
        set(w, L - 1);
        bad += check(d, v, L) + check(d, w, L);
        uint32_t all[4] = {0, 0, 0, 0}, one[4] = {0, 0, 0, 0};
        for (int i = 0; i < L; ++i) set(all, i);
        set(one, L / 2);
        bad += check(all, one, L);
        tot += 3;
    }
    std::printf("%ld %ld\n", bad, tot);
    return bad != 0;
}
"""


def _clang():
    for c in (os.environ.get("CXX_CLANG"), "/opt/rocm/llvm/bin/clang++", shutil.which("clang++")):
        if c and os.path.exists(c):
            return c
    return None


def test_row_fill_equals_site_closure(tmp_path):
    cxx = _clang()
    if cxx is None:
        pytest.skip("no clang++ for the host build of rowfill.hpp")
    src = tmp_path / "rowfill_check.cpp"
    exe = tmp_path / "rowfill_check"
    src.write_text(DRIVER)
    subprocess.run([cxx, "-std=c++17", "-O2", "-I", HDR_DIR, str(src), "-o", str(exe)], check=True,
                   capture_output=True, timeout=120)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    bad, tot = map(int, out.stdout.split())
    assert tot == 128 * 4003
    assert bad == 0 and out.returncode == 0


def test_row_fill_check_catches_a_dropped_wrap(tmp_path):
    """The check above is not vacuous: a fill without the wrap pass differs
    from the closure on rows whose run crosses the wrap."""
    cxx = _clang()
    if cxx is None:
        pytest.skip("no clang++ for the host build of rowfill.hpp")
    hdr = open(os.path.join(HDR_DIR, "rowfill.hpp")).read()
    broken = hdr.replace("if (!more) break;", "break;")
    assert broken != hdr
    (tmp_path / "rowfill.hpp").write_text(broken)
    src = tmp_path / "rowfill_check.cpp"
    exe = tmp_path / "rowfill_check"
    src.write_text(DRIVER)
    subprocess.run([cxx, "-std=c++17", "-O2", "-I", str(tmp_path), str(src), "-o", str(exe)], check=True,
                   capture_output=True, timeout=120)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    bad, _ = map(int, out.stdout.split())
    assert bad > 0
