"""The word-at-a-time UTF-8 helpers of the device code (utf8_offset,
utf8_chars, utf8_chars_rows in blazingsql_amd/csrc/vm_device.h), compiled for
the host with g++ and checked against a byte-by-byte loop (a byte counts as
a character unless it is a continuation byte 10xxxxxx, the rule of
cudf::strings::count_characters) over random mixes of ASCII and multi-byte
sequences at every start alignment and length, empty and absent strings
included. No GPU needed."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "blazingsql_amd", "csrc", "vm_device.h")

PRELUDE = r"""
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#define GLOBAL
#define __device__
#define __forceinline__ inline
static int __popc(uint32_t x) { return __builtin_popcount(x); }
static int __ffs(uint32_t x) { return __builtin_ffs(x); }
"""

MAIN = r"""
static bool is_lead(uint8_t c) { return (c & 0xC0) != 0x80; }
static int64_t bytes_chars(const uint8_t* s, int32_t n) {
  int64_t c = 0;
  for (int32_t i = 0; i < n; ++i) c += is_lead(s[i]);
  return c;
}
static int32_t bytes_offset(const uint8_t* s, int32_t n, int64_t chars) {
  if (n <= 0 || chars < 0) return n;
  int64_t c = 0;
  for (int32_t i = 0; i < n; ++i)
    if (is_lead(s[i])) {
      if (c == chars) return i;
      ++c;
    }
  return n;
}
int main() {
  alignas(16) static uint8_t buf[512];
  srand(5);
  long bad = 0;
  const uint8_t mix[6] = {0xC3, 0xA9, 0xE2, 0x82, 0xF0, 0x9F};
  for (int t = 0; t < 100000; ++t) {
    for (int i = 0; i < 512; ++i) buf[i] = rand() % 3 ? (uint8_t)('a' + rand() % 26) : mix[rand() % 6];
    const uint8_t* s[4];
    int32_t n[4];
    for (int k = 0; k < 4; ++k) {
      s[k] = buf + rand() % 128;
      n[k] = rand() % 80 - 4;
      if (rand() % 7 == 0) { s[k] = nullptr; n[k] = 0; }
    }
    if (s[0]) {
      const int64_t ch = rand() % 90 - 5;
      if (utf8_offset(s[0], n[0], ch) != bytes_offset(s[0], n[0], ch)) ++bad;
      if (utf8_chars(s[0], n[0]) != (n[0] > 0 ? bytes_chars(s[0], n[0]) : 0)) ++bad;
    }
    int64_t c4[4], c2[2], c1[1];
    utf8_chars_rows<4>(s, n, c4);
    const uint8_t* s2[2] = {s[1], s[2]};
    const int32_t n2[2] = {n[1], n[2]};
    utf8_chars_rows<2>(s2, n2, c2);
    const uint8_t* s1[1] = {s[3]};
    const int32_t n1[1] = {n[3]};
    utf8_chars_rows<1>(s1, n1, c1);
    for (int k = 0; k < 4; ++k)
      if (c4[k] != (n[k] > 0 ? bytes_chars(s[k], n[k]) : 0)) ++bad;
    for (int k = 0; k < 2; ++k)
      if (c2[k] != (n2[k] > 0 ? bytes_chars(s2[k], n2[k]) : 0)) ++bad;
    if (c1[0] != (n1[0] > 0 ? bytes_chars(s1[0], n1[0]) : 0)) ++bad;
  }
  printf("mismatches %ld\n", bad);
  return bad != 0;
}
"""


def _extract(src, start, end):
    i = src.index(start)
    j = src.index(end, i)
    return src[i:j]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_utf8_word_helpers_match_byte_loop(tmp_path):
    src = open(HEADER).read()
    offset = _extract(src, "static __device__ int32_t utf8_offset", "// Characters of the UTF-8 bytes")
    chars = _extract(src, "static __device__ int64_t utf8_chars", "// utf8_chars of a lane's R strings")
    rows = _extract(src, "template <int R>\n__device__ __forceinline__ void utf8_chars_rows",
                    "// The UTF-8 character (its encoded bytes")
    cpp = tmp_path / "utf8_words.cpp"
    cpp.write_text(PRELUDE + offset + chars + rows + MAIN)
    exe = tmp_path / "utf8_words"
    subprocess.run(["g++", "-O1", "-std=c++17", "-o", str(exe), str(cpp)], check=True, timeout=120)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
