/* C++ API of the engine: talsh::Tensor, talsh::TensorTask and the namespace
 * functions, the interface of TALSH/talshxx.hpp:24-620 and
 * TALSH/talsh_task.hpp (same names, signatures, default arguments and
 * argument meanings), implemented inline over the C-ABI of talsh.h. Header
 * only: include this file and link libtalsh.so.
 *
 * Differences a caller can observe:
 *  - errors come back as return codes (with a message on stderr); nothing
 *    aborts;
 *  - an operation asked to execute on DEV_HOST returns TALSH_NOT_AVAILABLE,
 *    as the C-ABI does: the engine computes on the GPU. Pass DEV_NVIDIA_GPU,
 *    DEV_AMD_GPU or DEV_DEFAULT as the execution device; tensors constructed
 *    here live on the host and are staged per the COPY_* rules;
 *  - pinHostMemory / unpinHostMemory leave the memory pageable and return
 *    TALSH_SUCCESS (transfers from it are staged, so results are unchanged);
 *  - orthogonalizeMGS returns TALSH_NOT_IMPLEMENTED (so does the
 *    reference's C-ABI); the SVD forms run on the GPU. */
#ifndef TALSHXX_HPP_
#define TALSHXX_HPP_

#include <complex>
#include <cstdio>
#include <initializer_list>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "talsh.h"
#include "talsh_task.hpp"

namespace talsh {

const std::size_t DEFAULT_HOST_BUFFER_SIZE = TALSH_NO_HOST_BUFFER;

const int REAL32 = R4;
const int REAL64 = R8;
const int COMPLEX32 = C4;
const int COMPLEX64 = C8;

/** Element type <-> TAL-SH data kind (TALSH/talshxx.hpp:37-80). **/
template <typename T>
struct TensorData {
  static constexpr int kind = NO_TYPE;
  static constexpr bool supported = false;
};
template <>
struct TensorData<float> {
  static constexpr int kind = R4;
  static constexpr bool supported = true;
  static constexpr float unity = 1.0f;
  static constexpr float zero = 0.0f;
};
template <>
struct TensorData<double> {
  static constexpr int kind = R8;
  static constexpr bool supported = true;
  static constexpr double unity = 1.0;
  static constexpr double zero = 0.0;
};
template <>
struct TensorData<std::complex<float>> {
  static constexpr int kind = C4;
  static constexpr bool supported = true;
  static constexpr std::complex<float> unity = {1.0f, 0.0f};
  static constexpr std::complex<float> zero = {0.0f, 0.0f};
};
template <>
struct TensorData<std::complex<double>> {
  static constexpr int kind = C8;
  static constexpr bool supported = true;
  static constexpr std::complex<double> unity = {1.0, 0.0};
  static constexpr std::complex<double> zero = {0.0, 0.0};
};

template <int K>
struct TensorDataType { using value = void; };
template <>
struct TensorDataType<R4> { using value = float; };
template <>
struct TensorDataType<R8> { using value = double; };
template <>
struct TensorDataType<C4> { using value = std::complex<float>; };
template <>
struct TensorDataType<C8> { using value = std::complex<double>; };

inline double realPart(float x) { return x; }
inline double realPart(double x) { return x; }
inline double realPart(std::complex<float> x) { return x.real(); }
inline double realPart(std::complex<double> x) { return x.real(); }
inline double imagPart(float) { return 0.0; }
inline double imagPart(double) { return 0.0; }
inline double imagPart(std::complex<float> x) { return x.imag(); }
inline double imagPart(std::complex<double> x) { return x.imag(); }

namespace detail {
inline int report(const char *where, const char *call, int errc)
{
  if (errc != TALSH_SUCCESS && errc != TRY_LATER && errc != DEVICE_UNABLE)
    std::fprintf(stderr, "#ERROR(talsh::%s): %s error %d\n", where, call, errc);
  return errc;
}
}  // namespace detail

/** Dense local tensor (TALSH/talshxx.hpp:97-556). Copies share one tensor
    block; the block is destroyed with the last copy, and a pending
    asynchronous task holds a copy until it completes. **/
class Tensor {
 public:
  /** Host-side typed view of a tensor slice (TALSH/talshxx.hpp:100-133):
      element [i0, i1, ...] of the view is element [bases + i] of the full
      tensor, dimension 0 fastest. **/
  template <typename FloatingType>
  class View {
   public:
    View(FloatingType *body, const std::vector<std::size_t> &full_extents, const std::vector<std::size_t> &extents,
         const std::vector<std::size_t> &bases)
        : body_(body), full_(full_extents), ext_(extents), base_(bases)
    {
    }

    template <typename IntegralType>
    FloatingType &operator[](const std::initializer_list<IntegralType> mlndx)
    {
      static_assert(std::is_integral<IntegralType>::value, "talsh::Tensor::View: the multi-index must be integral");
      std::size_t off = 0;
      const IntegralType *ix = mlndx.begin();
      for (std::size_t d = full_.size(); d-- > 0;) off = off * full_[d] + base_[d] + (std::size_t)ix[d];
      return body_[off];
    }

   private:
    FloatingType *body_;
    const std::vector<std::size_t> full_, ext_, base_;
  };

  template <typename T>
  Tensor(const std::initializer_list<std::size_t> signature, const std::initializer_list<int> dims, const T init_val)
      : pimpl_(std::make_shared<Impl>(std::vector<std::size_t>(signature), std::vector<int>(dims)))
  {
    pimpl_->construct_value(init_val);
  }
  template <typename T>
  Tensor(const std::vector<std::size_t> &signature, const std::vector<int> &dims, const T init_val)
      : pimpl_(std::make_shared<Impl>(signature, dims))
  {
    pimpl_->construct_value(init_val);
  }
  template <typename T>
  Tensor(const std::vector<std::size_t> &signature, const std::vector<int> &dims, const std::vector<T> &ext_data)
      : pimpl_(std::make_shared<Impl>(signature, dims))
  {
    pimpl_->construct_import(ext_data);
  }
  Tensor(const std::vector<std::size_t> &signature, const std::vector<int> &dims, int data_kind,
         talsh_tens_init_i init_func)
      : pimpl_(std::make_shared<Impl>(signature, dims))
  {
    pimpl_->construct_kind(data_kind, init_func);
  }
  template <typename T>
  Tensor(const std::initializer_list<std::size_t> signature, const std::initializer_list<int> dims, T *ext_mem,
         const T *init_val = nullptr)
      : pimpl_(std::make_shared<Impl>(std::vector<std::size_t>(signature), std::vector<int>(dims)))
  {
    pimpl_->construct_external(ext_mem, init_val);
  }
  template <typename T>
  Tensor(const std::vector<std::size_t> &signature, const std::vector<int> &dims, T *ext_mem,
         const T *init_val = nullptr)
      : pimpl_(std::make_shared<Impl>(signature, dims))
  {
    pimpl_->construct_external(ext_mem, init_val);
  }
  template <typename T>
  Tensor(const std::vector<int> &dims, const T init_val) : Tensor(std::vector<std::size_t>(dims.size(), 0), dims, init_val)
  {
  }
  template <typename T>
  Tensor(const std::vector<int> &dims, const std::vector<T> &ext_data)
      : Tensor(std::vector<std::size_t>(dims.size(), 0), dims, ext_data)
  {
  }
  Tensor(const std::vector<int> &dims, int data_kind, talsh_tens_init_i init_func)
      : Tensor(std::vector<std::size_t>(dims.size(), 0), dims, data_kind, init_func)
  {
  }
  template <typename T>
  Tensor(const std::vector<int> &dims, T *ext_mem, const T *init_val = nullptr)
      : Tensor(std::vector<std::size_t>(dims.size(), 0), dims, ext_mem, init_val)
  {
  }

  Tensor(const Tensor &) = default;
  Tensor &operator=(const Tensor &) = default;
  Tensor(Tensor &&) = default;
  Tensor &operator=(Tensor &&) = default;
  ~Tensor() = default;

  /** True if construction failed (e.g. no host memory). **/
  bool isEmpty() const { return talshTensorIsEmpty(&pimpl_->tensor_) != NOPE; }

  int getElementType() const
  {
    int n = 0, kinds[TALSH_MAX_DEV_PRESENT];
    if (talshTensorDataKind(&pimpl_->tensor_, &n, kinds) != TALSH_SUCCESS || n < 1) return NO_TYPE;
    return kinds[0];
  }
  int getRank() const { return talshTensorRank(&pimpl_->tensor_); }
  int getOrder() const { return getRank(); }
  std::size_t getVolume() const { return talshTensorVolume(&pimpl_->tensor_); }
  std::size_t getSize() const
  {
    int es = 0;
    return tens_valid_data_kind(getElementType(), &es) == YEP ? getVolume() * (std::size_t)es : 0;
  }

  const std::vector<std::size_t> &getDimOffsets() const { return pimpl_->signature_; }
  void resetDimOffsets(const std::vector<std::size_t> &offsets) { pimpl_->signature_ = offsets; }

  const int *getDimExtents(unsigned int &num_dims) const
  {
    int r = 0;
    const int *d = talshTensorDimExtents(&pimpl_->tensor_, &r);
    num_dims = (unsigned int)(r > 0 ? r : 0);
    return d;
  }
  int getDimExtent(unsigned int dim) const
  {
    unsigned int n = 0;
    const int *d = getDimExtents(n);
    return (d != nullptr && dim < n) ? d[dim] : -1;
  }

  /** New extents of the same volume (TALSH/talshxx.hpp:240-241). **/
  int reshape(const std::vector<int> &dims)
  {
    completeWriteTask();
    const int e = talshTensorReshape(&pimpl_->tensor_, (int)dims.size(), dims.data());
    if (e == TALSH_SUCCESS) pimpl_->signature_.assign(dims.size(), 0);
    return detail::report("Tensor::reshape", "talshTensorReshape", e);
  }

  /** Pointer to the host image, if the tensor has one of type T. **/
  template <typename T>
  bool getDataAccessHost(T **data_ptr)
  {
    completeWriteTask();
    void *p = nullptr;
    const int e = talshTensorGetBodyAccess(&pimpl_->tensor_, &p, TensorData<T>::kind, 0, DEV_HOST);
    *data_ptr = e == TALSH_SUCCESS ? static_cast<T *>(p) : nullptr;
    return e == TALSH_SUCCESS;
  }
  template <typename T>
  bool getDataAccessHostConst(const T **data_ptr)
  {
    completeWriteTask();
    const void *p = nullptr;
    const int e = talshTensorGetBodyAccessConst(&pimpl_->tensor_, &p, TensorData<T>::kind, 0, DEV_HOST);
    *data_ptr = e == TALSH_SUCCESS ? static_cast<const T *>(p) : nullptr;
    return e == TALSH_SUCCESS;
  }

  /** Views of the whole tensor or of a slice, on the host image (the
      tensor is first brought to the host exclusively). **/
  template <typename T>
  View<T> getSliceView()
  {
    std::vector<std::size_t> full = full_extents();
    return View<T>(host_body<T>(), full, full, std::vector<std::size_t>(full.size(), 0));
  }
  template <typename T, typename IntegralType>
  View<T> getSliceView(const std::initializer_list<IntegralType> slice_extents)
  {
    static_assert(std::is_integral<IntegralType>::value, "talsh::Tensor::getSliceView: extents must be integral");
    std::vector<std::size_t> full = full_extents();
    return View<T>(host_body<T>(), full, std::vector<std::size_t>(slice_extents.begin(), slice_extents.end()),
                   std::vector<std::size_t>(full.size(), 0));
  }
  template <typename T, typename IntegralType>
  View<T> getSliceView(const std::initializer_list<IntegralType> slice_extents,
                       const std::initializer_list<IntegralType> slice_bases)
  {
    static_assert(std::is_integral<IntegralType>::value, "talsh::Tensor::getSliceView: extents must be integral");
    std::vector<std::size_t> full = full_extents();
    return View<T>(host_body<T>(), full, std::vector<std::size_t>(slice_extents.begin(), slice_extents.end()),
                   std::vector<std::size_t>(slice_bases.begin(), slice_bases.end()));
  }

  /** Element access on the host image (slow path). **/
  template <typename T, typename IntegralType>
  T &operator[](const std::initializer_list<IntegralType> mlndx)
  {
    static_assert(std::is_integral<IntegralType>::value, "talsh::Tensor::operator[]: the multi-index must be integral");
    std::vector<std::size_t> full = full_extents();
    std::size_t off = 0;
    const IntegralType *ix = mlndx.begin();
    for (std::size_t d = full.size(); d-- > 0;) off = off * full[d] + (std::size_t)ix[d];
    return host_body<T>()[off];
  }

  Tensor &operator++()
  {
    ++pimpl_->used_;
    return *this;
  }
  Tensor &operator--()
  {
    if (pimpl_->used_ > 0) --pimpl_->used_;
    return *this;
  }

  /** Places the tensor on a device after its pending write task completed;
      exclusive discards the other images (TALSH/talshxx.hpp:273-289). **/
  bool sync(const int device_kind = DEV_HOST, const int device_id = 0, void *device_mem = nullptr,
            bool exclusive = false)
  {
    return sync(nullptr, device_kind, device_id, device_mem, exclusive);
  }
  bool sync(TensorTask *task_handle, const int device_kind = DEV_HOST, const int device_id = 0,
            void *device_mem = nullptr, bool exclusive = false)
  {
    if (!completeWriteTask()) return false;
    void *mem = device_mem != nullptr ? device_mem : (device_kind == DEV_HOST ? pimpl_->host_mem_ : nullptr);
    talsh_task_t *tk = begin_task(task_handle);
    const int e = detail::report("Tensor::sync", "talshTensorPlace",
                                 talshTensorPlace(&pimpl_->tensor_, device_id, device_kind, mem, COPY_M, tk));
    end_task(task_handle, e, {this});
    if (e != TALSH_SUCCESS) return false;
    if (exclusive) {
      if (task_handle != nullptr && !task_handle->wait()) return false;
      return talshTensorDiscardOther(&pimpl_->tensor_, device_id, device_kind) == TALSH_SUCCESS;
    }
    return true;
  }

  void discardImage(const int device_kind, const int device_id)
  {
    detail::report("Tensor::discardImage", "talshTensorDiscard",
                   talshTensorDiscard(&pimpl_->tensor_, device_id, device_kind));
  }

  /** True once the pending write task (if any) has completed; then the
      tensor is placed on the given device (TALSH/talshxx.hpp:291-297). **/
  bool ready(int *status, const int device_kind = DEV_HOST, const int device_id = 0, void *device_mem = nullptr)
  {
    *status = TALSH_TASK_EMPTY;
    if (!testWriteTask(status)) return false;
    if (*status == TALSH_TASK_COMPLETED || *status == TALSH_TASK_EMPTY) {
      void *mem = device_mem != nullptr ? device_mem : (device_kind == DEV_HOST ? pimpl_->host_mem_ : nullptr);
      detail::report("Tensor::ready", "talshTensorPlace",
                     talshTensorPlace(&pimpl_->tensor_, device_id, device_kind, mem, COPY_M, nullptr));
    }
    return true;
  }

  /** this = scalar_value **/
  template <typename T = double>
  int setValue(TensorTask *task_handle, const int device_kind = DEV_HOST, const int device_id = 0,
               const T scalar_value = TensorData<T>::zero)
  {
    completeWriteTask();
    talsh_task_t *tk = begin_task(task_handle);
    const int e = talshTensorInit(&pimpl_->tensor_, realPart(scalar_value), imagPart(scalar_value), device_id,
                                  device_kind, COPY_M, tk);
    end_task(task_handle, e, {this});
    return detail::report("Tensor::setValue", "talshTensorInit", e);
  }

  /** this *= scalar_value **/
  template <typename T = double>
  int scale(TensorTask *task_handle, const T scalar_value, const int device_kind = DEV_HOST, const int device_id = 0)
  {
    completeWriteTask();
    talsh_task_t *tk = begin_task(task_handle);
    const int e = talshTensorScale(&pimpl_->tensor_, realPart(scalar_value), imagPart(scalar_value), device_id,
                                   device_kind, COPY_M, tk);
    end_task(task_handle, e, {this});
    return detail::report("Tensor::scale", "talshTensorScale", e);
  }

  /** Sum of |elements| of the host image (the tensor is synced to the host
      first; the reference computes it there too, TALSH/talshxx.cpp:353-367). **/
  int norm1(TensorTask *task_handle, double *tens_norm1, const int device_kind = DEV_HOST, const int device_id = 0)
  {
    (void)task_handle; (void)device_kind; (void)device_id;
    if (!sync(DEV_HOST, 0)) return TALSH_FAILURE;
    *tens_norm1 = talshTensorImageNorm1_cpu(&pimpl_->tensor_);
    return TALSH_SUCCESS;
  }

  /** slice = this[offsets ...] (or +=) **/
  int extractSlice(TensorTask *task_handle, Tensor &slice, const std::vector<int> &offsets,
                   const int device_kind = DEV_HOST, const int device_id = 0, bool accumulative = false)
  {
    completeWriteTask();
    slice.completeWriteTask();
    talsh_task_t *tk = begin_task(task_handle);
    const int e = talshTensorSlice(&slice.pimpl_->tensor_, &pimpl_->tensor_, offsets.data(), device_id, device_kind,
                                   COPY_MT, accumulative ? YEP : NOPE, tk);
    end_task(task_handle, e, {&slice, this});
    return detail::report("Tensor::extractSlice", "talshTensorSlice", e);
  }

  /** this[offsets ...] = slice (or +=) **/
  int insertSlice(TensorTask *task_handle, Tensor &slice, const std::vector<int> &offsets,
                  const int device_kind = DEV_HOST, const int device_id = 0, bool accumulative = false)
  {
    completeWriteTask();
    slice.completeWriteTask();
    talsh_task_t *tk = begin_task(task_handle);
    const int e = talshTensorInsert(&pimpl_->tensor_, &slice.pimpl_->tensor_, offsets.data(), device_id, device_kind,
                                    COPY_MT, accumulative ? YEP : NOPE, tk);
    end_task(task_handle, e, {this, &slice});
    return detail::report("Tensor::insertSlice", "talshTensorInsert", e);
  }

  /** this = permuted left ("D(a,b,c)=L(c,a,b)") **/
  int copyBody(TensorTask *task_handle, const std::string &pattern, Tensor &left, const int device_kind = DEV_HOST,
               const int device_id = 0)
  {
    completeWriteTask();
    left.completeWriteTask();
    talsh_task_t *tk = begin_task(task_handle);
    const int e = talshTensorCopy(pattern.c_str(), &pimpl_->tensor_, &left.pimpl_->tensor_, device_id, device_kind,
                                  COPY_MT, tk);
    end_task(task_handle, e, {this, &left});
    return detail::report("Tensor::copyBody", "talshTensorCopy", e);
  }

  /** this += left * factor ("D(a,b)+=L(b,a)") **/
  template <typename T = double>
  int accumulate(TensorTask *task_handle, const std::string &pattern, Tensor &left, const int device_kind = DEV_HOST,
                 const int device_id = 0, const T factor = TensorData<T>::unity)
  {
    completeWriteTask();
    left.completeWriteTask();
    talsh_task_t *tk = begin_task(task_handle);
    const int e = talshTensorAdd(pattern.c_str(), &pimpl_->tensor_, &left.pimpl_->tensor_, realPart(factor),
                                 imagPart(factor), device_id, device_kind, COPY_MT, tk);
    end_task(task_handle, e, {this, &left});
    return detail::report("Tensor::accumulate", "talshTensorAdd", e);
  }

  /** this += left * right * factor (this = ... when not accumulative) **/
  template <typename T = double>
  int contractAccumulate(TensorTask *task_handle, const std::string &pattern, Tensor &left, Tensor &right,
                         const int device_kind = DEV_HOST, const int device_id = 0,
                         const T factor = TensorData<T>::unity, bool accumulative = true)
  {
    completeWriteTask();
    left.completeWriteTask();
    right.completeWriteTask();
    talsh_task_t *tk = begin_task(task_handle);
    const int e = talshTensorContract(pattern.c_str(), &pimpl_->tensor_, &left.pimpl_->tensor_,
                                      &right.pimpl_->tensor_, realPart(factor), imagPart(factor), device_id,
                                      device_kind, COPY_MTT, accumulative ? YEP : NOPE, tk);
    end_task(task_handle, e, {this, &left, &right});
    return detail::report("Tensor::contractAccumulate", "talshTensorContract", e);
  }

  /** The same for contractions larger than device memory; always blocking,
      the task handle (if any) is left empty (TALSH/talshxx.hpp:392-405). **/
  template <typename T = double>
  int contractAccumulateXL(TensorTask *task_handle, const std::string &pattern, Tensor &left, Tensor &right,
                           const int device_kind = DEV_HOST, const int device_id = 0,
                           const T factor = TensorData<T>::unity, bool accumulative = true)
  {
    (void)task_handle;
    completeWriteTask();
    left.completeWriteTask();
    right.completeWriteTask();
    return detail::report("Tensor::contractAccumulateXL", "talshTensorContractXL",
                          talshTensorContractXL(pattern.c_str(), &pimpl_->tensor_, &left.pimpl_->tensor_,
                                                &right.pimpl_->tensor_, realPart(factor), imagPart(factor),
                                                device_id, device_kind, accumulative ? YEP : NOPE));
  }

  /** Matrix-product form: the leading (rank(L) - nc) dimensions of L and the
      trailing (rank(R) - nc) dimensions of R are D's, in order, and the
      trailing nc of L are contracted with the leading nc of R, where
      nc = (rank(L) + rank(R) - rank(D)) / 2 (TALSH/talshxx.hpp:407-415). **/
  template <typename T = double>
  int multiplyAccumulate(TensorTask *task_handle, Tensor &left, Tensor &right, const int device_kind = DEV_HOST,
                         const int device_id = 0, const T factor = TensorData<T>::unity)
  {
    const int dr = getRank(), lr = left.getRank(), rr = right.getRank();
    if (lr + rr < dr || (lr + rr - dr) % 2 != 0) return TALSH_INVALID_ARGS;
    const int nc = (lr + rr - dr) / 2, nl = lr - nc, nr = rr - nc;
    if (nl < 0 || nr < 0) return TALSH_INVALID_ARGS;
    std::vector<int> dig;  // digital pattern: L's dimensions, then R's
    for (int i = 0; i < nl; ++i) dig.push_back(i + 1);
    for (int i = 0; i < nc; ++i) dig.push_back(-(i + 1));
    for (int i = 0; i < nc; ++i) dig.push_back(-(nl + 1 + i));
    for (int i = 0; i < nr; ++i) dig.push_back(nl + 1 + i);
    char sym[MAX_CONTRACTION_PATTERN_LEN];
    int len = 0, ierr = 0, conj = 0;
    get_contr_pattern_sym(&lr, &rr, &conj, dig.data(), sym, &len, &ierr);
    if (ierr != 0 || len <= 0 || len >= MAX_CONTRACTION_PATTERN_LEN) return TALSH_FAILURE;
    sym[len] = '\0';
    return contractAccumulate(task_handle, std::string(sym), left, right, device_kind, device_id, factor);
  }

  int decomposeSVD(TensorTask *task_handle, const std::string &pattern, Tensor &left, Tensor &right, Tensor &middle,
                   const int device_kind = DEV_HOST, const int device_id = 0)
  {
    (void)task_handle;
    return talshTensorDecomposeSVD(pattern.c_str(), &pimpl_->tensor_, &left.pimpl_->tensor_, &right.pimpl_->tensor_,
                                   &middle.pimpl_->tensor_, 'N', device_id, device_kind);
  }
  int decomposeSVDL(TensorTask *task_handle, const std::string &pattern, Tensor &left, Tensor &right,
                    const int device_kind = DEV_HOST, const int device_id = 0)
  {
    (void)task_handle;
    return talshTensorDecomposeSVDL(pattern.c_str(), &pimpl_->tensor_, &left.pimpl_->tensor_,
                                    &right.pimpl_->tensor_, device_id, device_kind);
  }
  int decomposeSVDR(TensorTask *task_handle, const std::string &pattern, Tensor &left, Tensor &right,
                    const int device_kind = DEV_HOST, const int device_id = 0)
  {
    (void)task_handle;
    return talshTensorDecomposeSVDR(pattern.c_str(), &pimpl_->tensor_, &left.pimpl_->tensor_,
                                    &right.pimpl_->tensor_, device_id, device_kind);
  }
  int decomposeSVDLR(TensorTask *task_handle, const std::string &pattern, Tensor &left, Tensor &right,
                     const int device_kind = DEV_HOST, const int device_id = 0)
  {
    (void)task_handle;
    return talshTensorDecomposeSVDLR(pattern.c_str(), &pimpl_->tensor_, &left.pimpl_->tensor_,
                                     &right.pimpl_->tensor_, device_id, device_kind);
  }
  int orthogonalizeSVD(TensorTask *task_handle, const std::string &pattern, const int device_kind = DEV_HOST,
                       const int device_id = 0)
  {
    (void)task_handle;
    return talshTensorOrthogonalizeSVD(pattern.c_str(), &pimpl_->tensor_, device_id, device_kind);
  }
  int orthogonalizeMGS(TensorTask *task_handle, const std::vector<unsigned int> &iso_dims,
                       const int device_kind = DEV_HOST, const int device_id = 0)
  {
    (void)task_handle;
    std::vector<int> dims(iso_dims.begin(), iso_dims.end());
    return talshTensorOrthogonalizeMGS(&pimpl_->tensor_, (int)dims.size(), dims.data(), device_id, device_kind);
  }

  void print() const { talshTensorPrintInfo(&pimpl_->tensor_); }
  void print(double thresh) const
  {
    talshTensorPrintInfo(&pimpl_->tensor_);
    talshTensorPrintBody(&pimpl_->tensor_, thresh);
  }

  /** The task that last wrote this tensor (non-owning), if still pending. **/
  void resetWriteTask(TensorTask *task = nullptr) { pimpl_->write_task_ = task; }
  TensorTask *getWriteTask() { return pimpl_->write_task_; }

  talsh_tens_t *getTalshTensorPtr() { return &pimpl_->tensor_; }

  friend int determineOptimalDevice(Tensor &tens0);
  friend int determineOptimalDevice(Tensor &tens0, Tensor &tens1);
  friend int determineOptimalDevice(Tensor &tens0, Tensor &tens1, Tensor &tens2);
  friend class TensorTask;

 private:
  struct Impl {
    std::vector<std::size_t> signature_;
    std::vector<int> dims_;
    talsh_tens_t tensor_;
    TensorTask *write_task_ = nullptr;
    void *host_mem_ = nullptr;  // application memory the host image lives in, if any
    int used_ = 0;

    Impl(const std::vector<std::size_t> &signature, const std::vector<int> &dims) : signature_(signature), dims_(dims)
    {
      talshTensorClean(&tensor_);
    }
    Impl(const Impl &) = delete;
    Impl &operator=(const Impl &) = delete;
    // a pending task holds a reference to this block, so none is pending here
    ~Impl() { talshTensorDestruct(&tensor_); }
    int host_dev() const { return talshFlatDevId(DEV_HOST, 0); }
    template <typename T>
    void construct_value(const T v)
    {
      static_assert(TensorData<T>::supported, "talsh::Tensor: unsupported element type");
      detail::report("Tensor::Tensor", "talshTensorConstruct",
                     talshTensorConstruct(&tensor_, TensorData<T>::kind, (int)dims_.size(), dims_.data(), host_dev(),
                                          nullptr, 0, nullptr, realPart(v), imagPart(v)));
    }
    template <typename T>
    void construct_import(const std::vector<T> &data)
    {
      static_assert(TensorData<T>::supported, "talsh::Tensor: unsupported element type");
      int e = talshTensorConstruct(&tensor_, TensorData<T>::kind, (int)dims_.size(), dims_.data(), host_dev(), nullptr,
                                   0, nullptr, 0.0, 0.0);
      if (e == TALSH_SUCCESS) {
        if (talshTensorVolume(&tensor_) > data.size()) e = TALSH_INVALID_ARGS;
        else e = talshTensorImportData(&tensor_, TensorData<T>::kind, static_cast<const void *>(data.data()));
      }
      detail::report("Tensor::Tensor", "talshTensorConstruct/ImportData", e);
    }
    void construct_kind(int data_kind, talsh_tens_init_i init_func)
    {
      detail::report("Tensor::Tensor", "talshTensorConstruct",
                     talshTensorConstruct(&tensor_, data_kind, (int)dims_.size(), dims_.data(), host_dev(), nullptr, 0,
                                          init_func, 0.0, 0.0));
    }
    template <typename T>
    void construct_external(T *mem, const T *init)
    {
      static_assert(TensorData<T>::supported, "talsh::Tensor: unsupported element type");
      host_mem_ = static_cast<void *>(mem);
      const int e = talshTensorConstruct(&tensor_, TensorData<T>::kind, (int)dims_.size(), dims_.data(), host_dev(),
                                         host_mem_, 0, nullptr, 0.0, 0.0);
      detail::report("Tensor::Tensor", "talshTensorConstruct", e);
      if (e == TALSH_SUCCESS && init != nullptr) {
        const std::size_t n = talshTensorVolume(&tensor_);
        for (std::size_t i = 0; i < n; ++i) mem[i] = *init;
      }
    }
  };

  std::shared_ptr<Impl> pimpl_;

  bool completeWriteTask()
  {
    TensorTask *t = pimpl_->write_task_;
    if (t == nullptr) return true;
    return t->wait();  // also unlinks this tensor
  }
  bool testWriteTask(int *status)
  {
    TensorTask *t = pimpl_->write_task_;
    if (t == nullptr) {
      *status = TALSH_TASK_EMPTY;
      return true;
    }
    return t->test(status);
  }
  std::vector<std::size_t> full_extents() const
  {
    unsigned int n = 0;
    const int *d = getDimExtents(n);
    std::vector<std::size_t> full(n);
    for (unsigned int i = 0; i < n; ++i) full[i] = (std::size_t)d[i];
    return full;
  }
  template <typename T>
  T *host_body()
  {
    sync(DEV_HOST, 0, pimpl_->host_mem_, true);
    T *p = nullptr;
    getDataAccessHost(&p);
    return p;
  }
  talsh_task_t *begin_task(TensorTask *task_handle)
  {
    if (task_handle == nullptr) return nullptr;
    if (!task_handle->isEmpty()) task_handle->clean();
    return task_handle->getTalshTaskPtr();
  }
  // after an asynchronous call: record the tensors (the first is the one
  // written) and link it to the task, or release the task on failure
  void end_task(TensorTask *task_handle, int errc, std::initializer_list<Tensor *> used)
  {
    if (task_handle == nullptr) return;
    if (errc != TALSH_SUCCESS) {
      task_handle->clean();
      return;
    }
    unsigned int k = 0;
    for (Tensor *t : used) {
      if (k == MAX_TENSOR_OPERANDS) break;
      task_handle->used_[k] = t;
      task_handle->keep_[k] = t->pimpl_;
      task_handle->slot_[k] = &t->pimpl_->write_task_;
      ++k;
    }
    task_handle->num_tensors_ = k;
    (*used.begin())->resetWriteTask(task_handle);
  }
};

// ---- namespace functions (TALSH/talshxx.hpp:558-620) ---------------------

/** Initialises the engine on every GPU (flat GPU ids 0..n-1). **/
inline int initialize(std::size_t *host_buffer_size = nullptr)
{
  int ngpu = 0;
  int e = talshDeviceCount(DEV_NVIDIA_GPU, &ngpu);
  if (e != TALSH_SUCCESS) return e;
  int gpus[MAX_GPUS_PER_NODE];
  if (ngpu > MAX_GPUS_PER_NODE) ngpu = MAX_GPUS_PER_NODE;
  for (int i = 0; i < ngpu; ++i) gpus[i] = i;
  int host_arg_max = 0;
  std::size_t dflt = DEFAULT_HOST_BUFFER_SIZE;
  e = talshInit(host_buffer_size != nullptr ? host_buffer_size : &dflt, &host_arg_max, ngpu, gpus, 0, nullptr, 0,
                nullptr);
  return detail::report("initialize", "talshInit", e);
}
inline int shutdown() { return talshShutdown(); }

/** No-op pinning: the memory stays pageable and transfers are staged. **/
template <typename T>
int pinHostMemory(T *host_ptr, std::size_t mem_size)
{
  (void)host_ptr; (void)mem_size;
  return TALSH_SUCCESS;
}
template <typename T>
int unpinHostMemory(T *host_ptr)
{
  (void)host_ptr;
  return TALSH_SUCCESS;
}

inline int getDeviceCount(int dev_kind)
{
  int n = 0;
  return talshDeviceCount(dev_kind, &n) == TALSH_SUCCESS ? n : 0;
}
inline std::size_t getDeviceMaxTensorSize(const int device_kind = DEV_HOST, const int device_id = 0)
{
  return talshDeviceTensorSize(device_id, device_kind);
}
inline std::size_t getDeviceMaxBufferSize(const int device_kind = DEV_HOST, const int device_id = 0)
{
  return talshDeviceBufferSize(device_id, device_kind);
}
inline void *getDeviceBufferBasePtr(const int device_kind = DEV_HOST, const int device_id = 0)
{
  return talshDeviceBufferBasePtr(device_id, device_kind);
}
inline double getTotalFlopCount() { return talshDeviceGetFlops(); }

inline int determineOptimalDevice(Tensor &tens0) { return talshDetermineOptimalDevice(&tens0.pimpl_->tensor_); }
inline int determineOptimalDevice(Tensor &tens0, Tensor &tens1)
{
  return talshDetermineOptimalDevice(&tens0.pimpl_->tensor_, &tens1.pimpl_->tensor_);
}
inline int determineOptimalDevice(Tensor &tens0, Tensor &tens1, Tensor &tens2)
{
  return talshDetermineOptimalDevice(&tens0.pimpl_->tensor_, &tens1.pimpl_->tensor_, &tens2.pimpl_->tensor_);
}

inline bool enableFastMath(int device_kind, int device_id = DEV_DEFAULT)
{
  return talshEnableFastMath(device_kind, device_id) == TALSH_SUCCESS;
}

inline void startMemManagerLog() { talshMemManagerLogStart(); }
inline void finishMemManagerLog() { talshMemManagerLogFinish(); }
inline void startTensorOpLog() { talshTensorOpLogStart(); }
inline void finishTensorOpLog() { talshTensorOpLogFinish(); }
inline void printStatistics() { talshStats(); }

}  // namespace talsh

#endif  // TALSHXX_HPP_
