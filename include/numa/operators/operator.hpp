#pragma once
// Operator contract of the matrix-free solver path, restated from the
// reference's include/numa/operators/operator.hpp:17-98 so that code written
// against it compiles unchanged against the MI355X backend (numa::hip):
//   * DVec                              operator.hpp:17
//   * NonlinearOperator                 operator.hpp:44-50
//   * SecondOrderOperator               operator.hpp:53-58
//   * OpCounters / Counted              operator.hpp:61-98
// The BLAS-1 helpers and the FD derivatives of that header are CPU code; the
// GPU versions are in the HIP library (include/numa_hip.h).

#include <concepts>
#include <cstddef>
#include <vector>

namespace numa::op {

using DVec = std::vector<double>;

// out = F(x), out = F'(x) v.  The callee sizes `out`.
template <class Op>
concept NonlinearOperator = requires(const Op& f, const DVec& x, const DVec& v, DVec& out) {
    { f.size() } -> std::convertible_to<std::size_t>;
    f.eval(x, out);
    f.jvp(x, v, out);
};

// adds out = F''(x)[v, v]
template <class Op>
concept SecondOrderOperator = NonlinearOperator<Op> &&
                              requires(const Op& f, const DVec& x, const DVec& v, DVec& out) {
                                  f.d2(x, v, out);
                              };

struct OpCounters {
    std::size_t n_F = 0;
    std::size_t n_JVP = 0;
    std::size_t n_D2 = 0;
};

// Counts calls without changing results.
template <class Op>
class Counted {
public:
    explicit Counted(const Op& op) : op_(op) {}
    std::size_t size() const { return op_.size(); }
    void eval(const DVec& x, DVec& out) const {
        ++c_.n_F;
        op_.eval(x, out);
    }
    void jvp(const DVec& x, const DVec& v, DVec& out) const {
        ++c_.n_JVP;
        op_.jvp(x, v, out);
    }
    void d2(const DVec& x, const DVec& v, DVec& out) const
        requires SecondOrderOperator<Op>
    {
        ++c_.n_D2;
        op_.d2(x, v, out);
    }
    const OpCounters& counters() const { return c_; }
    void reset() const { c_ = OpCounters{}; }
    // Additions for device solvers that call the wrapped operator's kernels
    // directly: the wrapped operator, and the calls the solver reports.
    const Op& base() const { return op_; }
    void add(const OpCounters& c) const {
        c_.n_F += c.n_F;
        c_.n_JVP += c.n_JVP;
        c_.n_D2 += c.n_D2;
    }

private:
    const Op& op_;
    mutable OpCounters c_;
};

}  // namespace numa::op
