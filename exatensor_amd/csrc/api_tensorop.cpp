// talsh.h tensor-operation objects (TALSH/talshc.cpp:2598-3300).
//
// The object lifecycle, argument/specification setters, the flop/byte
// accounting, the halving decomposition (used by talshTensorContractXL) and
// the staged execution of a contraction on argument slices: activate
// (construct one tensor per argument slice on the execution GPU), load input
// (slice L and R out of their tensors), execute (asynchronous contraction into
// the D slice), test, store output (accumulative insert of the D slice into its
// tensor), deactivate, and the progress state machine over those stages
// (TALSH/talshc.cpp:2708-3000). As in the reference, only contractions
// execute.
#include <chrono>
#include <vector>
#include <cstring>

#include "common.hpp"

extern "C" {

int talshTensorOpCreate(talsh_tens_op_t **op)
{
  if (!op) return TALSH_INVALID_ARGS;
  *op = (talsh_tens_op_t *)std::malloc(sizeof(talsh_tens_op_t));
  if (!*op) return TRY_LATER;
  return talshTensorOpClean(*op);
}

int talshTensorOpClean(talsh_tens_op_t *op)
{
  if (!op) return TALSH_INVALID_ARGS;
  op->opkind = TALSH_TENSOR_NOOP;
  op->data_kind = NO_TYPE;
  op->num_args = 0;
  for (int i = 0; i < MAX_TENSOR_OPERANDS; ++i) {
    talshTensorSliceClean(&op->tens_slice[i]);
    talshTensorClean(&op->tens_arg[i]);
  }
  op->symb_pattern = nullptr;
  op->alpha_real = 1.0;
  op->alpha_imag = 0.0;
  talshTaskClean(&op->task_handle);
  op->exec_dev_id = DEV_NULL;
  op->stage = TALSH_OP_EMPTY;
  op->time_started = op->time_scheduled = op->time_completed = op->time_finished = -1.0;
  return 0;
}

int talshTensorOpSetArgument(talsh_tens_op_t *op, const talsh_tens_t *tensor, const size_t *offsets, const int *dims)
{
  if (!op || !tensor) return TALSH_INVALID_ARGS;
  if (op->num_args >= MAX_TENSOR_OPERANDS) return TALSH_LIMIT_EXCEEDED;
  const int e = talshTensorSliceConstruct(&op->tens_slice[op->num_args], tensor, offsets, dims, nullptr, nullptr);
  if (e) return e;
  ++op->num_args;
  op->stage = TALSH_OP_PARTIAL;
  return 0;
}

int talshTensorOpSpecify(talsh_tens_op_t *op, int operation_kind, int data_kind, const char *symbolic_pattern,
                         double prefactor_real, double prefactor_imag)
{
  // the pattern string is referenced, not copied: the caller keeps it alive
  // for the life of the operation (as in TALSH/talshc.cpp:2665-2689)
  if (!op) return TALSH_INVALID_ARGS;
  if (tens_valid_data_kind(data_kind, nullptr) != YEP) return TALSH_INVALID_ARGS;
  if (op->opkind != TALSH_TENSOR_NOOP) return TALSH_NOT_ALLOWED;  // specified once
  op->opkind = operation_kind;
  op->data_kind = data_kind;
  op->symb_pattern = symbolic_pattern;
  op->alpha_real = prefactor_real;
  op->alpha_imag = prefactor_imag;
  op->stage = TALSH_OP_DEFINED;
  return 0;
}

int talshTensorOpSetExecDevice(talsh_tens_op_t *op, int dev_id, int dev_kind)
{
  if (!op || dev_id < 0) return TALSH_INVALID_ARGS;
  if (op->stage != TALSH_OP_DEFINED) return TALSH_NOT_ALLOWED;  // after Specify, before activation
  op->exec_dev_id = dev_kind == DEV_DEFAULT ? dev_id : talshFlatDevId(dev_kind, dev_id);
  return 0;
}

}  // extern "C"

namespace {

double wall_seconds()
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Flat id of the GPU that holds the argument slices: the preset execution
// device when it is a GPU, else GPU 0.
int slice_device(const talsh_tens_op_t *op)
{
  int kind = DEV_NULL;
  if (op->exec_dev_id != DEV_NULL && talshKindDevId(op->exec_dev_id, &kind) >= 0 && kind != DEV_HOST)
    return op->exec_dev_id;
  return talshFlatDevId(DEV_AMD_GPU, 0);
}

// Integer base offsets of an argument slice.
std::vector<int> slice_offsets(const talsh_tens_slice_t &sl)
{
  std::vector<int> offs(std::max(sl.shape.num_dim, 0));
  for (size_t j = 0; j < offs.size(); ++j) offs[j] = sl.bases.offsets ? (int)sl.bases.offsets[j] : 0;
  return offs;
}

}  // namespace

extern "C" {

int talshTensorOpActivate(talsh_tens_op_t *op)
{
  if (!op) return TALSH_INVALID_ARGS;
  if (op->opkind == TALSH_TENSOR_NOOP || op->stage != TALSH_OP_DEFINED) return TALSH_NOT_ALLOWED;
  op->time_started = wall_seconds();
  const int dev = slice_device(op);
  int e = 0;
  unsigned i = 0;
  for (; i < op->num_args && !e; ++i) {
    const talsh_tens_slice_t &sl = op->tens_slice[i];
    talsh_tens_t *t = &op->tens_arg[i];
    e = talshTensorClean(t);
    // slice bodies are created uninitialised: L and R are loaded, D is
    // overwritten by the contraction
    if (!e) e = talshTensorConstruct(t, op->data_kind, sl.shape.num_dim, sl.shape.dims, dev, nullptr, -1, nullptr,
                                     0.0, 0.0);
  }
  if (e) {  // leave the operation as it was
    for (unsigned j = 0; j < i; ++j)
      if (talshTensorIsEmpty(&op->tens_arg[j]) != YEP) talshTensorDestruct(&op->tens_arg[j]);
    return e;
  }
  op->stage = TALSH_OP_RESOURCED;
  return 0;
}

int talshTensorOpLoadInput(talsh_tens_op_t *op)
{
  if (!op) return TALSH_INVALID_ARGS;
  if (op->stage != TALSH_OP_RESOURCED) return TALSH_NOT_ALLOWED;
  const int dev = slice_device(op);
  for (unsigned i = 1; i < op->num_args; ++i) {  // input slices only
    talsh_tens_slice_t &sl = op->tens_slice[i];
    if (talshTensorRank(sl.tensor) != talshTensorRank(&op->tens_arg[i])) return TALSH_OBJECT_BROKEN;
    const std::vector<int> offs = slice_offsets(sl);
    const int e = talshTensorSlice(&op->tens_arg[i], sl.tensor, offs.data(), dev, DEV_DEFAULT, COPY_MT, NOPE, nullptr);
    if (e) return e;
  }
  op->stage = TALSH_OP_LOADED;
  return 0;
}

int talshTensorOpExecute(talsh_tens_op_t *op, int dev_id, int dev_kind)
{
  if (!op) return TALSH_INVALID_ARGS;
  if (op->stage != TALSH_OP_LOADED) return TALSH_NOT_ALLOWED;
  // the device: preset in the operation (then the call passes defaults), or
  // given here (and recorded)
  if (op->exec_dev_id != DEV_NULL) {
    if (dev_id != DEV_DEFAULT || dev_kind != DEV_DEFAULT) return TALSH_INVALID_ARGS;
    dev_id = talshKindDevId(op->exec_dev_id, &dev_kind);
  } else {
    if (dev_id == DEV_DEFAULT) return TALSH_INVALID_ARGS;
    if (dev_kind == DEV_DEFAULT) {
      op->exec_dev_id = dev_id;
      dev_id = talshKindDevId(op->exec_dev_id, &dev_kind);
    } else {
      op->exec_dev_id = talshFlatDevId(dev_kind, dev_id);
    }
  }
  if (op->opkind != TALSH_TENSOR_CONTRACT) return TALSH_NOT_IMPLEMENTED;
  if (op->num_args != 3 || !op->symb_pattern) return TALSH_INVALID_ARGS;
  op->time_scheduled = wall_seconds();
  const int e = talshTensorContract(op->symb_pattern, &op->tens_arg[0], &op->tens_arg[1], &op->tens_arg[2],
                                    op->alpha_real, op->alpha_imag, dev_id, dev_kind, COPY_TTT, NOPE,
                                    &op->task_handle);
  if (e) return e;
  op->stage = TALSH_OP_SCHEDULED;
  return 0;
}

int talshTensorOpTest(talsh_tens_op_t *op, int *completed, int wait)
{
  if (!op || !completed) return TALSH_INVALID_ARGS;
  *completed = NOPE;
  if (op->stage != TALSH_OP_SCHEDULED) return TALSH_NOT_ALLOWED;
  int sts = TALSH_TASK_EMPTY, e = 0;
  if (wait == YEP) {
    e = talshTaskWait(&op->task_handle, &sts);
    if (!e && sts == TALSH_TASK_COMPLETED) *completed = YEP;
  } else {
    const int ans = talshTaskComplete(&op->task_handle, &sts, &e);
    if (!e && ans == YEP && sts == TALSH_TASK_COMPLETED) *completed = YEP;
  }
  if (!e && sts == TALSH_TASK_ERROR) e = TALSH_TASK_ERROR;
  if (!e && *completed == YEP) {
    op->time_completed = wall_seconds();
    op->stage = TALSH_OP_COMPLETED;
  }
  return e;
}

int talshTensorOpStoreOutput(talsh_tens_op_t *op)
{
  if (!op) return TALSH_INVALID_ARGS;
  if (op->stage != TALSH_OP_COMPLETED) return TALSH_NOT_ALLOWED;
  if (op->num_args > 0) {
    talsh_tens_slice_t &sl = op->tens_slice[0];
    if (talshTensorRank(sl.tensor) != talshTensorRank(&op->tens_arg[0])) return TALSH_OBJECT_BROKEN;
    const std::vector<int> offs = slice_offsets(sl);
    const int e = talshTensorInsert(sl.tensor, &op->tens_arg[0], offs.data(), slice_device(op), DEV_DEFAULT, COPY_MT,
                                    YEP, nullptr);  // accumulative insert
    if (e) return e;
  }
  op->stage = TALSH_OP_STORED;
  return 0;
}

int talshTensorOpDeactivate(talsh_tens_op_t *op)
{
  if (!op) return TALSH_INVALID_ARGS;
  if (op->stage != TALSH_OP_RESOURCED && op->stage != TALSH_OP_STORED) return TALSH_NOT_ALLOWED;
  int ier = 0, e = 0;
  if (op->stage == TALSH_OP_STORED) ier = talshTaskDestruct(&op->task_handle);
  for (int i = (int)op->num_args - 1; i >= 0 && !e; --i)
    if (talshTensorIsEmpty(&op->tens_arg[i]) != YEP) e = talshTensorDestruct(&op->tens_arg[i]);
  op->time_finished = wall_seconds();
  if (e) return e;
  op->stage = op->stage == TALSH_OP_STORED ? TALSH_OP_RETIRED : TALSH_OP_DEFINED;
  return ier ? NOT_CLEAN : 0;
}

int talshTensorOpProgress(talsh_tens_op_t *op, int *done)
{
  if (!done) return TALSH_INVALID_ARGS;
  *done = NOPE;
  if (!op) return TALSH_INVALID_ARGS;
  // synchronous stages run back to back; scheduling the contraction and a
  // negative completion test yield to the caller
  for (;;) {
    int e = 0, completed = NOPE;
    switch (op->stage) {
    case TALSH_OP_DEFINED: e = talshTensorOpActivate(op); break;
    case TALSH_OP_RESOURCED: e = talshTensorOpLoadInput(op); break;
    case TALSH_OP_LOADED: return talshTensorOpExecute(op, DEV_DEFAULT, DEV_DEFAULT);
    case TALSH_OP_SCHEDULED:
      e = talshTensorOpTest(op, &completed, NOPE);
      if (e || completed != YEP) return e;
      break;
    case TALSH_OP_COMPLETED: e = talshTensorOpStoreOutput(op); break;
    case TALSH_OP_STORED: e = talshTensorOpDeactivate(op); break;
    case TALSH_OP_RETIRED: *done = YEP; return 0;
    default: return TALSH_NOT_ALLOWED;
    }
    if (e) return e;
  }
}
// Splits a contraction in two along one dimension, halving it (the first half
// gets the odd element): a batch index (in L, R and D) if any is larger than 1,
// else whichever of the right (R,D), left (L,D) and contracted (L,R) index
// groups has the largest volume, taking that group's largest extent (the last
// one on ties). This is the rule of TALSH/talshc.cpp:3150-3294; the children
// carry the parent's pattern and prefactor.
int talshTensorOpDecompose2(const talsh_tens_op_t *op, talsh_tens_op_t *c1, talsh_tens_op_t *c2)
{
  if (!op || !c1 || !c2) return TALSH_INVALID_ARGS;
  if (op->opkind != TALSH_TENSOR_CONTRACT) return TALSH_NOT_IMPLEMENTED;
  if (!op->symb_pattern || op->num_args != 3) return TALSH_INVALID_ARGS;
  int dig[2 * MAX_TENSOR_RANK], dr, lr, rr, cj;
  int e = talsh_get_contr_ptrn_str2dig(op->symb_pattern, dig, &dr, &lr, &rr, &cj);
  if (e) return e;
  if (dr <= 0 && lr <= 0 && rr <= 0) return TALSH_NOT_ALLOWED;
  if (lr + rr <= 0) return TALSH_NOT_ALLOWED;
  const int *ldims = op->tens_slice[1].shape.dims, *rdims = op->tens_slice[2].shape.dims;
  std::vector<int> dst(dr > 0 ? dr : 1, 0);  // D position -> 1 + L position
  size_t lc = 1, ll = 1, lrv = 1, lb = 1, sc = 0, sl = 0, sr = 0, sb = 0;
  int mc = -1, ml = -1, mr = -1, mb = -1;
  for (int i = 0; i < lr; ++i) {
    const size_t cd = (size_t)ldims[i];
    if (dig[i] < 0) {
      lc *= cd;
      if (cd >= sc) { sc = cd; mc = i; }
    } else if (dig[i] > 0) {
      dst[dig[i] - 1] = i + 1;
      ll *= cd;
      if (cd >= sl) { sl = cd; ml = i; }
    }
  }
  for (int j = 0; j < rr; ++j) {
    const size_t cd = (size_t)rdims[j];
    if (dig[lr + j] > 0) {
      if (dst[dig[lr + j] - 1] == 0) {
        lrv *= cd;
        if (cd >= sr) { sr = cd; mr = lr + j; }
      } else {  // batch index
        ll /= cd;
        lb *= cd;
        if (cd >= sb) { sb = cd; mb = lr + j; }
      }
    }
  }
  if (!(lb > 1 || lrv > 1 || ll > 1 || lc > 1)) return TALSH_INVALID_REQUEST;
  int d0 = -1, d1 = -1, d2 = -1;  // dimension to halve in D, L, R (-1: none)
  if (lb > 1) {
    d0 = dig[mb] - 1; d1 = dst[d0] - 1; d2 = mb - lr;
  } else if (lrv >= ll && lrv >= lc) {
    d0 = dig[mr] - 1; d2 = mr - lr;
  } else if (ll >= lrv && ll >= lc) {
    d0 = dig[ml] - 1; d1 = ml;
  } else {
    d1 = mc; d2 = -dig[mc] - 1;
  }
  const int split[3] = {d0, d1, d2};
  for (int a = 0; a < 3 && !e; ++a) {
    const talsh_tens_slice_t &sl = op->tens_slice[a];
    const int rank = sl.shape.num_dim;
    std::vector<size_t> offs(rank > 0 ? rank : 1);
    std::vector<int> dims(rank > 0 ? rank : 1);
    for (int i = 0; i < rank; ++i) { offs[i] = sl.bases.offsets[i]; dims[i] = sl.shape.dims[i]; }
    const int d = split[a];
    if (d >= 0) {
      const int h1 = (dims[d] + 1) / 2, h2 = dims[d] - h1;
      dims[d] = h1;
      e = talshTensorOpSetArgument(c1, sl.tensor, offs.data(), dims.data());
      dims[d] = h2;
      offs[d] += (size_t)h1;
      if (!e) e = talshTensorOpSetArgument(c2, sl.tensor, offs.data(), dims.data());
    } else {
      e = talshTensorOpSetArgument(c1, sl.tensor, offs.data(), dims.data());
      if (!e) e = talshTensorOpSetArgument(c2, sl.tensor, offs.data(), dims.data());
    }
  }
  if (!e) e = talshTensorOpSpecify(c1, op->opkind, op->data_kind, op->symb_pattern, op->alpha_real, op->alpha_imag);
  if (!e) e = talshTensorOpSpecify(c2, op->opkind, op->data_kind, op->symb_pattern, op->alpha_real, op->alpha_imag);
  return e;
}

int talshTensorOpDestruct(talsh_tens_op_t *op)
{
  if (!op) return TALSH_INVALID_ARGS;
  if (talshTaskIsEmpty(&op->task_handle) != YEP) {
    // an operation whose contraction still runs cannot be destructed
    const int e = talshTaskDestruct(&op->task_handle);
    if (e) return e;
  }
  for (int i = 0; i < MAX_TENSOR_OPERANDS; ++i) {
    talshTensorSliceDestruct(&op->tens_slice[i]);
    if (talshTensorIsEmpty(&op->tens_arg[i]) != YEP) talshTensorDestruct(&op->tens_arg[i]);
  }
  return talshTensorOpClean(op);
}

int talshTensorOpDestroy(talsh_tens_op_t *op)
{
  if (!op) return TALSH_INVALID_ARGS;
  const int e = talshTensorOpDestruct(op);
  std::free(op);
  return e;
}

size_t talshTensorOpGetArgVolume(const talsh_tens_op_t *op, unsigned int arg_num)
{
  if (!op || arg_num >= op->num_args) return 0;
  return talshTensorSliceVolume(&op->tens_slice[arg_num]);
}

size_t talshTensorOpGetArgSize(const talsh_tens_op_t *op, unsigned int arg_num)
{
  int es = 0;
  if (!op || tens_valid_data_kind(op->data_kind, &es) != YEP) return 0;
  return talshTensorOpGetArgVolume(op, arg_num) * (size_t)es;
}

double talshTensorOpGetByteCount(const talsh_tens_op_t *op, unsigned int element_size)
{
  if (!op || op->opkind == TALSH_TENSOR_NOOP) return 0.0;
  int es = (int)element_size;
  if (op->data_kind != NO_TYPE && tens_valid_data_kind(op->data_kind, &es) != YEP) return 0.0;
  double b = 0;
  for (unsigned i = 0; i < op->num_args; ++i) b += (double)talshTensorSliceVolume(&op->tens_slice[i]) * es;
  return b;
}

double talshTensorOpGetFlopCount(const talsh_tens_op_t *op)
{
  // flops = fma * (product of distinct index extents), TALSH/talshc.cpp:3090-3139
  if (!op) return 0.0;
  double fma;
  if (op->data_kind == R4 || op->data_kind == R8) fma = 2.0;
  else if (op->data_kind == C4 || op->data_kind == C8) fma = 8.0;
  else return 0.0;
  if (op->opkind != TALSH_TENSOR_CONTRACT || !op->symb_pattern || op->num_args < 3) return 0.0;
  int dig[2 * MAX_TENSOR_RANK], dr, lr, rr, cj;
  if (talsh_get_contr_ptrn_str2dig(op->symb_pattern, dig, &dr, &lr, &rr, &cj)) return 0.0;
  std::vector<int> seen(dr + 1, 0);
  double f = 1.0;
  for (int i = 0; i < lr; ++i) {
    const double nd = op->tens_slice[1].shape.dims[i];
    if (dig[i] > 0) {
      if (!seen[dig[i]]) { seen[dig[i]] = 1; f *= nd; }
    } else {
      f *= nd;
    }
  }
  for (int j = 0; j < rr; ++j) {
    const double nd = op->tens_slice[2].shape.dims[j];
    const int v = dig[lr + j];
    if (v > 0 && !seen[v]) { seen[v] = 1; f *= nd; }
  }
  return f * fma;
}

double talshTensorOpGetIntensity(const talsh_tens_op_t *op)
{
  const double f = talshTensorOpGetFlopCount(op), b = talshTensorOpGetByteCount(op, 1);
  if (b <= 0.0 || f < 0.0) return -1.0;
  return f / b;
}

void talshTensorOpPrint(const talsh_tens_op_t *op)
{
  if (!op) return;
  std::printf("#MESSAGE: Tensor operation: kind %d, data kind %d, args %u, stage %d, pattern %s\n", op->opkind,
              op->data_kind, op->num_args, op->stage, op->symb_pattern ? op->symb_pattern : "(none)");
}

}  // extern "C"
