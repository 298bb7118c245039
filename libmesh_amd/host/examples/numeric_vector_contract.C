// numeric_vector_contract.C -- the NumericVector contract checks of the
// reference's tests/numerics/numeric_vector_test.h (Localize, LocalizeToOne,
// LocalizeIndices, Norms, Operations, WriteAndRead; each through the base
// class and the derived class) run on HipVector<Number>: one rank, block
// size 10, so global size = local size = 10.  Prints one line per check and
// "N failures".
#include <cmath>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "libmesh/libmesh.h"
#include "libmesh/numeric_vector.h"

using namespace libMesh;

static int failures = 0;
static void check(bool ok, const std::string &what) {
  std::printf("%s %s\n", ok ? "ok  " : "FAIL", what.c_str());
  if (!ok) ++failures;
}
// LIBMESH_ASSERT_NUMBERS_EQUAL / LIBMESH_ASSERT_FP_EQUAL with TOLERANCE^2
static bool near(Real a, Real b, Real tol = TOLERANCE * TOLERANCE) { return std::abs(a - b) <= tol; }

template <class Base, class Derived>
struct Suite {
  const Parallel::Communicator &comm;
  const unsigned block_size = 10, local_size = 10, global_size = 10;
  std::string name;

  void Operations() {
    auto v_ptr = std::make_unique<Derived>(comm, global_size, local_size);
    Base &v = *v_ptr;
    const dof_id_type first = v.first_local_index(), last = v.last_local_index();
    for (dof_id_type n = first; n != last; n++) v.set(n, static_cast<Number>(n + 1));
    v.close();
    auto v_clone = v.clone();
    auto &vorig = *v_clone;
    check(relative_fuzzy_equals(v, vorig), name + " Operations: clone equals");
    v += v;
    check(!relative_fuzzy_equals(v, vorig), name + " Operations: v += v differs");
    bool ok = true;
    for (dof_id_type n = first; n != last; n++) ok &= near(v(n), Real(2 * n + 2));
    check(ok, name + " Operations: v += v");
    v *= v;
    ok = true;
    for (dof_id_type n = first; n != last; n++) ok &= near(v(n), Real(4 * n * n + 8 * n + 4));
    check(ok, name + " Operations: v *= v");
    v -= vorig;
    ok = true;
    for (dof_id_type n = first; n != last; n++) ok &= near(v(n), Real(4 * n * n + 7 * n + 3));
    check(ok, name + " Operations: v -= vorig");
    v /= vorig;
    ok = true;
    for (dof_id_type n = first; n != last; n++) ok &= near(v(n), Real(4 * n + 3));
    check(ok, name + " Operations: v /= vorig");
    v.add(1);
    ok = true;
    for (dof_id_type n = first; n != last; n++) ok &= near(v(n), Real(4 * n + 4));
    check(ok, name + " Operations: add(1)");
    v.add(2, vorig);
    ok = true;
    for (dof_id_type n = first; n != last; n++) ok &= near(v(n), Real(6 * n + 6));
    check(ok, name + " Operations: add(2, vorig)");
    v.scale(1 / Real(3));
    ok = true;
    for (dof_id_type n = first; n != last; n++) ok &= near(v(n), Real(2 * n + 2));
    check(ok, name + " Operations: scale(1/3)");
    v = 4;
    ok = true;
    for (dof_id_type n = first; n != last; n++) ok &= near(v(n), Real(4));
    check(ok, name + " Operations: v = 4");
    v = vorig;
    ok = true;
    for (dof_id_type n = first; n != last; n++) ok &= near(v(n), Real(n + 1));
    check(ok, name + " Operations: v = vorig");
    v.reciprocal();
    ok = true;
    for (dof_id_type n = first; n != last; n++) ok &= near(v(n), 1 / Real(n + 1));
    check(ok, name + " Operations: reciprocal");
    check(near(v.dot(vorig), Real(global_size)), name + " Operations: dot");
  }

  void WriteAndRead() {
    auto v_ptr = std::make_unique<Derived>(comm, global_size, local_size);
    Base &v = *v_ptr;
    for (dof_id_type n = v.first_local_index(); n != v.last_local_index(); n++) v.set(n, static_cast<Number>(n + 1));
    v.close();
    v.print_matlab(name + "_vector.m");
    auto v_new_ptr = std::make_unique<Derived>(comm, global_size, local_size);
    Base &v_new = *v_new_ptr;
    v_new.read_matlab(name + "_vector.m");
    check(v.l1_norm_diff(v_new) == Real(0), name + " WriteAndRead: l1_norm_diff == 0");
  }

  void Norms() {
    auto v_ptr = std::make_unique<Derived>(comm, global_size, local_size);
    Base &v = *v_ptr;
    const dof_id_type first = v.first_local_index(), last = v.last_local_index();
    for (dof_id_type n = first; n != last; n++) v.set(n, -static_cast<Number>(n));
    v.close();
    for (dof_id_type n = first; n != last; n++) v.add(n, -static_cast<Number>(n));
    v.close();
    const Real exact_l1 = global_size * (global_size - Real(1));
    const Real exact_l2 = std::sqrt(Real(global_size - 1) * (2 * global_size) * (2 * global_size - 1) / 3);
    check(near(v.sum(), -exact_l1), name + " Norms: sum");
    check(near(v.l1_norm(), exact_l1), name + " Norms: l1_norm");
    check(near(v.l2_norm(), exact_l2), name + " Norms: l2_norm");
    check(near(v.linfty_norm(), 2 * (global_size - Real(1))), name + " Norms: linfty_norm");
    auto u_ptr = std::make_unique<Derived>(comm, global_size, local_size);
    Base &u = *u_ptr;
    for (dof_id_type n = first; n != last; n++) u.set(n, -static_cast<Number>(n * n));
    u.close();
    Derived diff_derived(comm, global_size, local_size);
    Base &diff = diff_derived;
    diff = u;
    diff -= v;
    diff.close();
    const auto diff_norm = diff.l2_norm();
    const auto norm_diff = u.l2_norm_diff(v);
    check(near(diff_norm, norm_diff, (std::abs(diff_norm) + std::abs(norm_diff)) * TOLERANCE * TOLERANCE),
          name + " Norms: l2_norm_diff");
  }

  void Localize(bool to_one) {
    const char *which = to_one ? " LocalizeToOne" : " Localize";
    auto v_ptr = std::make_unique<Derived>(comm, global_size, local_size);
    Base &v = *v_ptr;
    std::vector<Number> l(global_size);
    const dof_id_type first = v.first_local_index(), last = v.last_local_index();
    for (dof_id_type n = first; n != last; n++) v.set(n, static_cast<Number>(n));
    v.close();
    for (dof_id_type n = first; n != last; n++) v.add(n, static_cast<Number>(n));
    v.close();
    if (!to_one)
      v.localize(l);
    else
      v.localize_to_one(l, 0);
    bool ok = true;
    for (dof_id_type i = 0; i < v.size(); i++) ok &= near(2. * i, l[i]);
    check(ok, name + which + ": set + add");
    for (dof_id_type n = first; n != last; n++) {
      const auto value = static_cast<Number>(n);
      v.insert(&value, std::vector<numeric_index_type>({n}));
    }
    v.close();
    if (!to_one)
      v.localize(l);
    else
      v.localize_to_one(l, 0);
    ok = true;
    for (dof_id_type i = 0; i < v.size(); i++) ok &= near(Real(i), l[i]);
    check(ok, name + which + ": insert");
  }

  void LocalizeIndices() {
    auto v_ptr = std::make_unique<Derived>(comm, global_size, local_size);
    Base &v = *v_ptr;
    std::vector<Number> values(block_size);
    std::vector<numeric_index_type> indices;
    for (dof_id_type n = v.first_local_index(); n != v.last_local_index(); n++) v.set(n, static_cast<Number>(n));
    v.close();
    const dof_id_type end_index = block_size;
    for (unsigned int j = 0; j != block_size; ++j) indices.push_back(end_index - j - 1);
    v.localize(values, indices);
    bool ok = true;
    for (unsigned int j = 0; j != block_size; ++j) ok &= near(values[j], Real(end_index - j - 1));
    check(ok, name + " LocalizeIndices");
  }

  void run() {
    Localize(false);
    Localize(true);
    LocalizeIndices();
    Norms();
    Operations();
    WriteAndRead();
  }
};

int main(int argc, char **argv) {
  LibMeshInit init(argc, argv);
  try {
    Suite<HipVector<Number>, HipVector<Number>>{init.comm(), 10, 10, 10, "HipVector"}.run();
    Suite<NumericVector<Number>, HipVector<Number>>{init.comm(), 10, 10, 10, "HipVectorBase"}.run();
    // NumericVector::build for the GPU package
    auto b = NumericVector<Number>::build(init.comm());
    const numeric_index_type ten = 10;
    b->init(ten, ten);
    *b = 2.0;
    check(b->sum() == 20.0 && b->size() == 10, "NumericVector::build(HIP_SOLVERS)");
  } catch (const std::exception &e) {
    std::printf("FAIL exception: %s\n", e.what());
    return 1;
  }
  std::printf("%d failures\n", failures);
  return failures ? 1 : 0;
}
