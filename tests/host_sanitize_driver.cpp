// Host-layer sanitizer driver (test infrastructure; built by
// tests/test_sanitizers.py with AddressSanitizer + UndefinedBehaviorSanitizer
// on the host side of every library source).  Exercises, without a GPU, the
// host code that runs before any kernel launch: handle creation and
// teardown, argument validation of every entry point, workspace sizing, the
// ragged schedule builder (sorting, size classes, tiles, offsets) on valid
// and invalid offset arrays, and the tuning-cache parser.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../include/tokura_blas.h"

#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            std::exit(1);                                          \
        }                                                          \
    } while (0)

int main(int argc, char **argv)
{
    tokurablas_t *h = nullptr;
    const int rc = tokuraCreate(&h);
    CHECK(rc == 0 || rc == 1);   // 1: no HIP device here
    CHECK(h != nullptr);
    hipDoubleComplex *p = reinterpret_cast<hipDoubleComplex *>(0x1000);
    char *f = reinterpret_cast<char *>(0x2000);
    // validation happens before any device access
    CHECK(tokura_zgeev_batched_gpu(nullptr, 4, 1, p, p, p, f, nullptr) == -1);
    CHECK(tokura_zgeev_batched_gpu(h, 65, 1, p, p, p, f, nullptr) == -2);
    CHECK(tokura_zgeev_batched_gpu(h, 4, 1, nullptr, p, p, f, nullptr) == -3);
    CHECK(tokura_zgeev_batched_gpu(h, 4, 0, p, p, p, f, nullptr) == -6);
    CHECK(tokura_zgeev_batched_vectors_gpu(h, 4, 1, p, p, nullptr, nullptr, nullptr, f, nullptr) == -5);
    CHECK(tokura_dgeev_batched_gpu(h, 4, 1, nullptr, nullptr, nullptr, nullptr, f, nullptr) == -3);
    CHECK(tokura_get_zgeeveigenvaluesgetworspacesize(64, 100000) == 6656000000ull);
    CHECK(tokura_get_zgeeveigenvaluesgetworspacesize(0, 1) == (size_t)-1);
    CHECK(tokura_get_dgeev_worksize(64, 2) == 8u * (63 * 66 / 2 + 64) * 2);
    CHECK(tokura_get_zgeev_vectors_worksize(3, 2) == 16u * 9 * 2);
    CHECK(tokuraSetBalancing(h, 'B') == 0 && tokuraSetBalancing(h, 'x') == -2);
    CHECK(tokuraSetMaxSweeps(h, 5) == 0 && tokuraSetMaxSweeps(h, -1) == 0);
    CHECK(tokuraSetVariant(h, 64, 20) == 0);
    CHECK(tokuraSetVariant(h, 64, 1000) != 0);
    for (int c = 0; c < 4; ++c) CHECK(tokuraSetRaggedConfig(h, c) == 0);
    CHECK(tokuraSetRaggedConfig(h, 99) != 0);
    int cfg, l1, l2, l3, l4;
    CHECK(tokuraGetRaggedConfig(h, &cfg, &l1, &l2, &l3, &l4) == 0);
    // ragged schedule: random sizes 1..64, then malformed offset arrays
    std::srand(3);
    for (int trial = 0; trial < 20; ++trial) {
        const int batch = 1 + std::rand() % 3000;
        std::vector<int64_t> off(batch + 1, 0);
        for (int b = 0; b < batch; ++b) {
            const int n = 1 + std::rand() % 64;
            off[b + 1] = off[b] + (int64_t)n * n;
        }
        const size_t ws = tokura_get_zgeev_ragged_worksize(batch, off.data());
        CHECK(ws > 0 && ws < ((size_t)1 << 40));
        CHECK(tokura_get_zgeev_ragged_schur_worksize(batch, off.data()) > 0);
    }
    std::vector<int64_t> bad = {0, 16, 16 + 15};   // second count 15: not a square
    CHECK(tokura_get_zgeev_ragged_worksize(2, bad.data()) == (size_t)-3);
    std::vector<int64_t> big = {0, 65 * 65};       // n = 65
    CHECK(tokura_get_zgeev_ragged_worksize(1, big.data()) == (size_t)-3);
    std::vector<int64_t> neg = {0, -4};
    CHECK(tokura_get_zgeev_ragged_worksize(1, neg.data()) == (size_t)-3);
    CHECK(tokura_get_zgeev_ragged_worksize(0, bad.data()) == (size_t)-2);
    // tuning cache parser: a file with good, foreign and garbage lines
    if (argc > 1) {
        FILE *fp = std::fopen(argv[1], "w");
        CHECK(fp != nullptr);
        std::fprintf(fp, "# arch cus n variant\ngfx950 256 64 20\ngfx942 304 64 18\n");
        std::fprintf(fp, "garbage line without numbers\ngfx950 256 999 20\ngfx950 256 8 -7\n\n");
        std::fclose(fp);
        const int got = tokuraLoadTuning(h, argv[1]);
        CHECK(got >= -1 && got <= 2);
        CHECK(tokuraLoadTuning(h, "/nonexistent/tune.txt") < 0);
    }
    CHECK(tokuraDestroy(h) == 0);
    CHECK(tokuraDestroy(nullptr) == 0);
    std::printf("sanitized host run: no report\n");
    return 0;
}
