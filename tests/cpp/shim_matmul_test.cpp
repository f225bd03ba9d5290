// The reference's C++ matmul unit tests (test/unit/test_grg.cpp:41-161), run against the
// device-resident graph through grgl_hip.hpp.  Exit code 0 = all checks passed.
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "grgl_hip.hpp"

using grgl_hip::DeviceGRG;

static int failures = 0;
#define EXPECT(cond)                                                                                  \
    do {                                                                                              \
        if (!(cond)) {                                                                                \
            std::printf("FAILED: %s (line %d)\n", #cond, __LINE__);                                  \
            failures++;                                                                               \
        }                                                                                             \
    } while (0)

int main() {
    // depth3BinTree (test/unit/common_grgs.h:13-26): samples 0-3, 4->(0,1), 5->(2,3), 6->(4,5)
    const std::vector<uint64_t> off = {0, 0, 0, 0, 0, 2, 4, 6};
    const std::vector<uint32_t> kids = {0, 1, 2, 3, 4, 5};
    DeviceGRG grg(4, 2, off, kids, /*mutation nodes=*/{1, 5, 6});
    EXPECT(grg.numNodes() == 7 && grg.numMutations() == 3);

    // DotProductGood
    std::vector<double> mutValues = {1.0, 2.0, 1.0};
    EXPECT((grg.matMul(mutValues, 1, grgl_hip::DIRECTION_DOWN) == std::vector<double>{1.0, 2.0, 3.0, 3.0}));
    std::vector<double> sampleValues(4, 1.0);
    EXPECT((grg.matMul(sampleValues, 1, grgl_hip::DIRECTION_UP) == std::vector<double>{1.0, 2.0, 4.0}));

    // MatrixMultGood
    std::vector<double> mv(6, 1.0);
    mv[1] = 2.0;
    mv[3] = 0.0;
    EXPECT((grg.matMul(mv, 2, grgl_hip::DIRECTION_DOWN) ==
            std::vector<double>{1.0, 2.0, 3.0, 3.0, 1.0, 1.0, 2.0, 2.0}));
    std::vector<double> sv(8, 1.0);
    for (size_t i = 4; i < 8; i++) sv[i] /= 4.0;
    EXPECT((grg.matMul(sv, 2, grgl_hip::DIRECTION_UP) == std::vector<double>{1.0, 2.0, 4.0, 0.25, 0.5, 1.0}));

    // Integer input, accumulated into the caller's output (the reference's `+=`).
    std::vector<int64_t> in = {11, 41, 37, 73};
    std::vector<int64_t> out = {100, 100, 100};
    grg.matrixMultiplication<int64_t>(in.data(), 4, 1, grgl_hip::DIRECTION_UP, out.data(), 3);
    EXPECT((out == std::vector<int64_t>{141, 210, 262}));

    // MatMulBad
    int thrown = 0;
    try { grg.matMul(std::vector<double>(4), 1, grgl_hip::DIRECTION_DOWN); } catch (const grgl_hip::ApiMisuseFailure&) { thrown++; }
    try { grg.matMul(std::vector<double>(3), 1, grgl_hip::DIRECTION_UP); } catch (const grgl_hip::ApiMisuseFailure&) { thrown++; }
    try { grg.matMul(std::vector<double>(9), 2, grgl_hip::DIRECTION_UP); } catch (const grgl_hip::ApiMisuseFailure&) { thrown++; }
    try { grg.matMul(std::vector<double>(9), 9, grgl_hip::DIRECTION_UP); } catch (const grgl_hip::ApiMisuseFailure&) { thrown++; }
    EXPECT(thrown == 4);

    // storeGPUGRGToDisk / loadGPUGRGFromDisk: the reloaded graph multiplies the same.
    const char* path = std::getenv("GRGL_SHIM_LVLZ");
    if (path) {
        grg.store(path);
        std::unique_ptr<grgl_hip::DeviceGRG> again = grgl_hip::DeviceGRG::load(path);
        EXPECT(again->numSamples() == 4 && again->numMutations() == 3 && again->numNodes() == 7);
        EXPECT((again->matMul(sv, 2, grgl_hip::DIRECTION_UP) == std::vector<double>{1.0, 2.0, 4.0, 0.25, 0.5, 1.0}));
        EXPECT((again->matMul(mv, 2, grgl_hip::DIRECTION_DOWN) ==
                std::vector<double>{1.0, 2.0, 3.0, 3.0, 1.0, 1.0, 2.0, 2.0}));
    }

    std::printf("%s (%d failures)\n", failures ? "FAIL" : "OK", failures);
    return failures ? 1 : 0;
}
