// Checks of the C++ API (include/talshxx.hpp) through the engine on GPU 0:
// the closed-form contraction checks of the reference's C++ test
// (TALSH/test.cpp:166-293: first contraction with beta 0, second one
// accumulating, matrix-multiplication form, slice extraction / value reset /
// insertion) run on the GPU instead of the host, plus asynchronous task
// polling, complex data, permuted copy, accumulate, norm1 and the XL form.
// Prints one line per check; exits non-zero if any check fails.
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "talshxx.hpp"

namespace {

int failures = 0;

void check(bool ok, const char *what, double got, double want)
{
  std::printf("%s %s: got %.12g want %.12g\n", ok ? "PASS" : "FAIL", what, got, want);
  if (!ok) ++failures;
}

bool close(double a, double b, double rel) { return std::fabs(a - b) <= rel * std::fmax(std::fabs(b), 1e-300); }

}  // namespace

int main()
{
  const int VDIM = 40, ODIM = 20;
  const int dev = DEV_NVIDIA_GPU;  // the MI355X, as existing callers name it
  std::size_t host_buf = std::size_t(1) << 30;
  if (talsh::initialize(&host_buf) != TALSH_SUCCESS) {
    std::printf("FAIL initialize\n");
    return 1;
  }
  check(talsh::getDeviceCount(DEV_NVIDIA_GPU) >= 1, "GPU count", talsh::getDeviceCount(DEV_NVIDIA_GPU), 1);
  check(talsh::getDeviceMaxBufferSize(dev, 0) > 0, "device buffer size > 0", (double)talsh::getDeviceMaxBufferSize(dev, 0), 1);
  {
    // contraction, beta 0 then accumulating (TALSH/test.cpp:192-230)
    talsh::Tensor dtens({1, 2, 3, 4}, {VDIM, VDIM, ODIM, ODIM}, 1.0);
    talsh::Tensor ltens({5, 6, 7, 8}, {ODIM, VDIM, ODIM, VDIM}, 0.01);
    talsh::Tensor rtens({9, 10, 11, 12}, {VDIM, VDIM, VDIM, VDIM}, 0.001);
    check(!dtens.isEmpty() && dtens.getRank() == 4 && dtens.getVolume() == (std::size_t)VDIM * VDIM * ODIM * ODIM,
          "construct rank/volume", (double)dtens.getVolume(), (double)VDIM * VDIM * ODIM * ODIM);
    check(dtens.getElementType() == talsh::REAL64, "element type", dtens.getElementType(), talsh::REAL64);
    talsh::TensorTask task;
    int e = dtens.contractAccumulate(&task, "D(a,b,c,d)+=L(d,i,c,j)*R(j,b,i,a)", ltens, rtens, dev, 0, 0.5, false);
    check(e == TALSH_SUCCESS, "contractAccumulate (async) returns success", e, TALSH_SUCCESS);
    check(dtens.getWriteTask() == &task, "write task recorded", 1, 1);
    check(dtens.sync(), "sync after contraction", 1, 1);
    check(dtens.getWriteTask() == nullptr && task.isEmpty(), "sync completes and empties the task", 1, 1);
    const double *dp = nullptr;
    dtens.getDataAccessHostConst(&dp);
    const double v1 = 0.01 * 0.001 * VDIM * VDIM * 0.5;
    bool all = dp != nullptr;
    for (std::size_t i = 0; all && i < dtens.getVolume(); ++i) all = close(dp[i], v1, 1e-12);
    check(all, "first contraction, every element", dp ? dp[0] : NAN, v1);
    task.clean();
    e = dtens.contractAccumulate(&task, "D(a,b,c,d)+=L(d,i,c,j)*R(j,b,i,a)", ltens, rtens, dev, 0, 0.5, true);
    check(e == TALSH_SUCCESS && dtens.sync(), "second contraction", e, TALSH_SUCCESS);
    dtens.getDataAccessHostConst(&dp);
    check(dp != nullptr && close(dp[0], 2 * v1, 1e-12), "second contraction accumulates", dp ? dp[0] : NAN, 2 * v1);
    auto view = dtens.getSliceView<double>();
    check(close(view[{VDIM - 1, 3, ODIM - 1, 7}], 2 * v1, 1e-12), "view element", view[{VDIM - 1, 3, ODIM - 1, 7}], 2 * v1);
    double n1 = 0;
    e = dtens.norm1(nullptr, &n1);
    check(e == TALSH_SUCCESS && close(n1, 2 * v1 * (double)dtens.getVolume(), 1e-9), "norm1", n1,
          2 * v1 * (double)dtens.getVolume());
    // the same contraction through the XL form (blocking)
    talsh::Tensor xtens({VDIM, VDIM, ODIM, ODIM}, 0.0);
    e = xtens.contractAccumulateXL(nullptr, "D(a,b,c,d)+=L(d,i,c,j)*R(j,b,i,a)", ltens, rtens, dev, 0, 0.5, true);
    const double *xp = nullptr;
    xtens.sync();
    xtens.getDataAccessHostConst(&xp);
    check(e == TALSH_SUCCESS && xp != nullptr && close(xp[123], v1, 1e-12), "contractAccumulateXL", xp ? xp[123] : NAN, v1);
  }
  {
    // matrix-multiplication form (TALSH/test.cpp:233-247)
    talsh::Tensor dtens({1, 2, 3, 4}, {VDIM, VDIM, ODIM, ODIM}, 0.0);
    talsh::Tensor ltens({5, 6, 7, 8}, {VDIM, VDIM, VDIM, VDIM}, 0.01);
    talsh::Tensor rtens({9, 10, 11, 12}, {VDIM, VDIM, ODIM, ODIM}, 0.001);
    talsh::TensorTask task;
    int e = dtens.multiplyAccumulate(&task, ltens, rtens, dev, 0, 0.5);
    check(e == TALSH_SUCCESS && dtens.sync(), "multiplyAccumulate", e, TALSH_SUCCESS);
    const double *dp = nullptr;
    dtens.getDataAccessHostConst(&dp);
    const double want = 0.5 * 0.01 * 0.001 * VDIM * VDIM;
    check(dp != nullptr && close(dp[dtens.getVolume() - 1], want, 1e-12), "multiplyAccumulate value",
          dp ? dp[dtens.getVolume() - 1] : NAN, want);
  }
  {
    // slice extraction, value reset, insertion (TALSH/test.cpp:250-287)
    talsh::Tensor ltens({0, 0, 0}, {3, 4, 5}, 0.0);
    unsigned int nd = 0;
    const int *ld = ltens.getDimExtents(nd);
    double *lp = nullptr;
    ltens.getDataAccessHost(&lp);
    for (int i = 0; i < ld[0] * ld[1] * ld[2]; ++i) lp[i] = i;
    talsh::Tensor dtens({1, 2, 3}, {2, 2, 2}, 0.0);
    talsh::TensorTask task;
    int e = ltens.extractSlice(&task, dtens, std::vector<int>{1, 2, 3}, dev, 0);
    check(e == TALSH_SUCCESS && dtens.sync(), "extractSlice", e, TALSH_SUCCESS);
    auto sv = dtens.getSliceView<double>();
    bool ok = true;
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) ok = ok && sv[{i, j, k}] == (double)((1 + i) + (2 + j) * 3 + (3 + k) * 12);
    check(ok, "slice elements", sv[{1, 1, 1}], (double)(2 + 3 * 3 + 4 * 12));
    task.clean();
    e = dtens.setValue(&task, dev, 0, -13.0);
    check(e == TALSH_SUCCESS && dtens.sync(), "setValue", e, TALSH_SUCCESS);
    task.clean();
    e = ltens.insertSlice(&task, dtens, std::vector<int>{1, 2, 3}, dev, 0);
    check(e == TALSH_SUCCESS && ltens.sync(), "insertSlice", e, TALSH_SUCCESS);
    ltens.getDataAccessHost(&lp);
    int wrong = 0;
    for (int k = 0; k < 5; ++k)
      for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 3; ++i) {
          const bool in = i >= 1 && i < 3 && j >= 2 && j < 4 && k >= 3 && k < 5;
          const int l = i + 3 * j + 12 * k;
          wrong += lp[l] != (in ? -13.0 : (double)l);
        }
    check(wrong == 0, "insertSlice elements (wrong count)", wrong, 0);
  }
  {
    // asynchronous completion by polling, complex data
    typedef std::complex<float> cf;
    talsh::Tensor dtens({ODIM, ODIM}, cf{0.5f, 0.0f});
    talsh::Tensor ltens({ODIM, VDIM, ODIM, VDIM}, cf{0.01f, 0.02f});
    talsh::Tensor rtens({VDIM, VDIM}, cf{0.001f, -0.003f});
    talsh::TensorTask task;
    int e = dtens.contractAccumulate(&task, "D(a,c)+=L(a,i,c,j)*R(j,i)", ltens, rtens, dev, 0, cf{1.0f, 0.0f});
    int st = TALSH_TASK_EMPTY, polls = 0;
    while (e == TALSH_SUCCESS && !task.test(&st)) ++polls;
    check(e == TALSH_SUCCESS && st == TALSH_TASK_COMPLETED, "complex contraction completes by polling", st,
          TALSH_TASK_COMPLETED);
    dtens.sync();
    const cf *cp = nullptr;
    dtens.getDataAccessHostConst(&cp);
    const std::complex<double> want = std::complex<double>(0.5, 0.0) +
                                      std::complex<double>(0.01f, 0.02f) * std::complex<double>(0.001f, -0.003f) *
                                          (double)(VDIM * VDIM);
    check(cp != nullptr && close(cp[17].real(), want.real(), 1e-5), "complex contraction value (real part)",
          cp ? cp[17].real() : NAN, want.real());
    check(cp != nullptr && close(cp[17].imag(), want.imag(), 1e-4), "complex contraction value (imaginary part)",
          cp ? cp[17].imag() : NAN, want.imag());
  }
  {
    // imported data, permuted copy and accumulate
    std::vector<double> a(6 * 7);
    for (std::size_t i = 0; i < a.size(); ++i) a[i] = 0.5 * (double)i;
    talsh::Tensor ltens(std::vector<int>{6, 7}, a);
    talsh::Tensor dtens(std::vector<int>{7, 6}, 0.0);
    int e = dtens.copyBody(nullptr, "D(a,b)=L(b,a)", ltens, dev, 0);
    dtens.sync();
    double *dp = nullptr;
    dtens.getDataAccessHost(&dp);
    int wrong = 0;
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 7; ++j) wrong += dp == nullptr || dp[j + 7 * i] != a[i + 6 * j];
    check(e == TALSH_SUCCESS && wrong == 0, "copyBody transpose (wrong count)", wrong, 0);
    e = dtens.accumulate(nullptr, "D(a,b)+=L(b,a)", ltens, dev, 0, 2.0);
    dtens.sync();
    dtens.getDataAccessHost(&dp);
    check(e == TALSH_SUCCESS && dp != nullptr && dp[3 + 7 * 2] == 3.0 * a[2 + 6 * 3], "accumulate",
          dp ? dp[3 + 7 * 2] : NAN, 3.0 * a[2 + 6 * 3]);
    e = dtens.scale(nullptr, 0.25, dev, 0);
    dtens.sync();
    dtens.getDataAccessHost(&dp);
    check(e == TALSH_SUCCESS && dp != nullptr && dp[5] == 0.75 * a[6 * 5], "scale", dp ? dp[5] : NAN, 0.75 * a[6 * 5]);
    // operations on the host are not served by the engine
    e = dtens.scale(nullptr, 2.0, DEV_HOST, 0);
    check(e == TALSH_NOT_AVAILABLE, "DEV_HOST execution is refused", e, TALSH_NOT_AVAILABLE);
  }
  talsh::shutdown();
  std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "ALL PASSED", failures);
  return failures ? 1 : 0;
}
