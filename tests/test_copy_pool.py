"""The host threads behind large search results and query packing (csrc/copypool.hpp), built
alone with g++ and run on the CPU: every byte of copies of odd and chunk-sized lengths arrives,
run() gives every part exactly once, pools are created and joined repeatedly, and one pool
serves many jobs in a row."""

from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

CSRC = Path(__file__).resolve().parents[1] / "hnsw-engine_amd" / "csrc"

PROGRAM = r"""
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include "copypool.hpp"

int main() {
    const size_t sizes[] = {0, 1, 63, 64, 65, 255, 4096, 100003, (size_t)4 << 20,
                            ((size_t)4 << 20) + 3, 9999991};
    std::vector<unsigned char> src(((size_t)4 << 20) + 4096 + 9999991), dst(src.size());
    for (size_t i = 0; i < src.size(); ++i) src[i] = (unsigned char)(i * 2654435761u >> 13);
    for (int workers = 0; workers <= 5; ++workers) {
        hm::CopyPool pool(workers);
        for (int rep = 0; rep < 3; ++rep) {
            for (size_t n : sizes) {
                for (size_t off : {(size_t)0, (size_t)7}) {
                    memset(dst.data(), 0xAB, n + off + 16);
                    pool.copy(dst.data() + off, src.data() + off, n);
                    if (memcmp(dst.data() + off, src.data() + off, n) != 0) {
                        printf("mismatch workers=%d n=%zu off=%zu\n", workers, n, off);
                        return 1;
                    }
                    for (size_t i = n + off; i < n + off + 16; ++i)
                        if (dst[i] != 0xAB) {
                            printf("overrun workers=%d n=%zu off=%zu\n", workers, n, off);
                            return 1;
                        }
                }
            }
        }
    }
    {  // run(): every part once, on its own index range
        hm::CopyPool pool(3);
        for (int rep = 0; rep < 100; ++rep) {
            std::vector<int> hit(1000, 0);
            pool.run([&](size_t p, size_t parts) {
                for (size_t i = p; i < hit.size(); i += parts) hit[i] += 1;
            });
            for (int h : hit)
                if (h != 1) {
                    puts("run missed or repeated an index");
                    return 1;
                }
        }
    }
    for (int i = 0; i < 200; ++i) {  // create and join repeatedly
        hm::CopyPool pool(3);
        pool.copy(dst.data(), src.data(), 1000);
    }
    puts("ok");
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_copy_pool_copies_every_byte(tmp_path) -> None:
    src = tmp_path / "pool_test.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "pool_test"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", f"-I{CSRC}", str(src), "-o", str(exe)],
                   check=True, capture_output=True, text=True, timeout=120)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=120)
    assert out.stdout.strip() == "ok"
