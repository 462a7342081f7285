// Text driver around the reference's indexed max-heap (inc/heap.hpp), used
// only to pin the oracle's heap restatement (tests/test_oracle_heap.py).
// stdin:  n v_0 .. v_{n-1}   then lines "u key value" (update) or "p" (peek)
// stdout: one "key value" line per peek, value printed with %.17g
#include <cstdio>
#include <vector>
#include "heap.hpp"

int main() {
    size_t n = 0;
    if (std::scanf("%zu", &n) != 1 || n == 0) return 2;
    std::vector<double> values(n);
    for (size_t i = 0; i < n; ++i) {
        if (std::scanf("%lf", &values[i]) != 1) return 2;
    }
    Heap<double> heap(values);
    char op[4];
    while (std::scanf("%3s", op) == 1) {
        if (op[0] == 'u') {
            size_t key; double value;
            if (std::scanf("%zu %lf", &key, &value) != 2) return 2;
            heap.update(key, value);
        } else if (op[0] == 'p') {
            HeapItem<double> item = heap.peek();
            std::printf("%zu %.17g\n", item.key, item.value);
        }
    }
    return 0;
}
