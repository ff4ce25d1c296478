/* C++ task handle of the engine: talsh::TensorTask, the interface of
 * TALSH/talsh_task.hpp:21-75 (same members and meanings), implemented inline
 * over the C-ABI of talsh.h. Part of the header-only C++ API (talshxx.hpp).
 *
 * A TensorTask owns one talsh_task_t. An asynchronous talsh::Tensor
 * operation fills an empty task and records the tensors it uses, holding a
 * reference to their tensor blocks so that none is destroyed while the
 * device still works on it; wait() and a successful test() complete the task
 * and return it to the empty state, releasing those references and the
 * tensors' write-task links. Errors are reported by return values and a
 * message on stderr, never by aborting. */
#ifndef TALSH_TASK_HPP_
#define TALSH_TASK_HPP_

#include <cstdio>
#include <memory>

#include "talsh.h"

namespace talsh {

class Tensor;

class TensorTask {
 public:
  TensorTask() { talshTaskClean(&task_); }
  TensorTask(const TensorTask &) = delete;
  TensorTask &operator=(const TensorTask &) = delete;
  TensorTask(TensorTask &&) = delete;
  TensorTask &operator=(TensorTask &&) = delete;
  ~TensorTask() { wait(); }

  /** True when no TAL-SH task is attached (TALSH/talsh_task.hpp:33-34). **/
  bool isEmpty() { return talshTaskIsEmpty(&task_) == YEP; }

  /** Completes an active task if needed, unlinks the tensors it wrote and
      returns the handle to the empty state (TALSH/talsh_task.hpp:36-37). **/
  void clean()
  {
    if (!isEmpty()) {
      const int st = talshTaskStatus(&task_);
      if (st != TALSH_TASK_ERROR && st != TALSH_TASK_EMPTY && st != TALSH_TASK_COMPLETED) {
        int stats = 0;
        talshTaskWait(&task_, &stats);
      }
    }
    for (unsigned int i = 0; i < num_tensors_; ++i) {
      if (slot_[i] != nullptr && *slot_[i] == this) *slot_[i] = nullptr;
      slot_[i] = nullptr;
      used_[i] = nullptr;
      keep_[i].reset();  // may release the last reference to a tensor block
    }
    num_tensors_ = 0;
    talshTaskDestruct(&task_);
    talshTaskClean(&task_);
  }

  /** Completes the task; true if it completed without error
      (TALSH/talsh_task.hpp:39-40). An empty task counts as completed. **/
  bool wait()
  {
    int stats = TALSH_TASK_COMPLETED;
    if (!isEmpty()) {
      int ierr = TALSH_SUCCESS;
      if (talshTaskComplete(&task_, &stats, &ierr) != YEP) {
        const int e = talshTaskWait(&task_, &stats);
        if (e != TALSH_SUCCESS) std::fprintf(stderr, "#ERROR(talsh::TensorTask::wait): talshTaskWait error %d\n", e);
      }
      if (stats != TALSH_TASK_COMPLETED)
        std::fprintf(stderr, "#ERROR(talsh::TensorTask::wait): task completed with status %d\n", stats);
    }
    clean();
    return stats == TALSH_TASK_COMPLETED;
  }

  /** Tests for completion: true (and the handle cleaned) once the task has
      completed, status set to its TALSH_TASK_* state; an empty task reports
      TALSH_TASK_EMPTY and true (TALSH/talsh_task.hpp:42-43). **/
  bool test(int *status)
  {
    if (isEmpty()) {
      *status = TALSH_TASK_EMPTY;
      clean();
      return true;
    }
    int ierr = TALSH_SUCCESS;
    const bool done = talshTaskComplete(&task_, status, &ierr) == YEP;
    if (ierr != TALSH_SUCCESS)
      std::fprintf(stderr, "#ERROR(talsh::TensorTask::test): talshTaskComplete error %d\n", ierr);
    if (done) clean();
    return done;
  }

  /** Execution device id (kind in *device_kind); negative if undefined. **/
  int getExecutionDevice(int *device_kind) { return talshTaskDevId(&task_, device_kind); }

  unsigned int getNumTensorArguments() const { return num_tensors_; }

  const Tensor *getTensorArgument(unsigned int arg_num) const
  {
    return arg_num < num_tensors_ ? used_[arg_num] : nullptr;
  }

  const talshTensArg_t *getTensorArgumentImages(int *num_arguments)
  {
    return talshTaskTensArgs(&task_, num_arguments);
  }

  int getTensorArgumentCoherence() { return talshTaskArgCoherence(&task_); }

  talsh_task_t *getTalshTaskPtr() { return &task_; }

 private:
  talsh_task_t task_;
  unsigned int num_tensors_ = 0;
  Tensor *used_[MAX_TENSOR_OPERANDS] = {};             // the handles the operation was called with
  std::shared_ptr<void> keep_[MAX_TENSOR_OPERANDS];    // their tensor blocks, kept alive until completion
  TensorTask **slot_[MAX_TENSOR_OPERANDS] = {};        // their write-task links

  friend class Tensor;
};

}  // namespace talsh

#endif  // TALSH_TASK_HPP_
