"""The refinement queue's heap pops in std::priority_queue's order.

The adaptive refinement (finite_source_magnifier.cpp:3201-3283) splits the
cell at the top of a std::priority_queue ordered by error weight; cells of
equal weight (common: a uniform source's weights are 0 or 1) leave the queue in
the order libstdc++'s push_heap / pop_heap give them, and that order decides
which cells are split.  lcbinint_amd/csrc/pq_heap.hpp restates those two heap
walks; this test compiles it on the host with g++ beside a
std::priority_queue and checks that both return the same payloads in the same
order over random push/pop sequences with many ties.
"""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]

DRIVER = r"""
#include <cstdio>
#include <queue>
#include <random>
#include <vector>
#include "pq_heap.hpp"

struct Item {
    double key;
    int val;
    bool operator<(const Item& o) const { return key < o.key; }
};

int main()
{
    std::mt19937_64 rng(7);
    long checked = 0;
    for (int trial = 0; trial < 400; ++trial) {
        const int levels = 1 + static_cast<int>(rng() % 6);  // few distinct keys: many ties
        std::vector<double> keys(80000);
        std::vector<int> vals(80000);
        lcbg::PqHeap<> h{keys.data(), vals.data(), 0, {}};
        std::priority_queue<Item> ref;
        int next = 0;
        const int ops = 200 + static_cast<int>(rng() % 3000);
        for (int op = 0; op < ops; ++op) {
            const bool push = ref.empty() || rng() % 5 < 3;
            if (push) {
                const double k = trial % 3 == 0 ? static_cast<double>(rng() % 1000) * 0.25
                                                : static_cast<double>(rng() % levels) * 0.5;
                h.push(k, next);
                ref.push({k, next});
                ++next;
            } else {
                double k = 0.0;
                const int v = h.pop(&k);
                const Item t = ref.top();
                ref.pop();
                if (v != t.val || k != t.key) {
                    std::printf("MISMATCH trial %d op %d: %d/%g vs %d/%g\n", trial, op, v, k, t.val, t.key);
                    return 1;
                }
                ++checked;
            }
            if (h.n != static_cast<int>(ref.size())) {
                std::printf("SIZE trial %d op %d\n", trial, op);
                return 1;
            }
        }
        while (!ref.empty()) {
            double k = 0.0;
            const int v = h.pop(&k);
            if (v != ref.top().val) {
                std::printf("MISMATCH drain trial %d\n", trial);
                return 1;
            }
            ref.pop();
            ++checked;
        }
    }
    std::printf("OK %ld\n", checked);
    return 0;
}
"""


def test_heap_pops_in_priority_queue_order(tmp_path):
    src = tmp_path / "heap_order.cpp"
    exe = tmp_path / "heap_order"
    src.write_text(DRIVER)
    subprocess.run(["g++", "-O1", "-std=c++17", f"-I{ROOT / 'lcbinint_amd' / 'csrc'}", str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("OK"), out.stdout
    assert int(out.stdout.split()[1]) > 100000
