// batch.cpp — many independent instances at once (no reference
// counterpart: the reference builds one GLR per call).  Instances small
// enough that one build leaves most of the GPU idle (short kernels, and
// host decisions between them: Cholesky and eigensolver results, interval
// estimates) are spread over host worker threads; each worker drives the
// engine on its own stream (device.hip), so the kernels of several builds
// overlap on the device and the host work (AMG set-up, launches) runs in
// parallel.  Instances that share an object (a prior, B, Z or mass) form
// one group, handled in order by one worker, since calls on one object
// must not overlap.
#include "engine.hpp"
#include "petsc_impl.hpp"

#include <algorithm>
#include <atomic>
#include <functional>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

using namespace lgh;

namespace {

// groups of instance indices that share any of the listed objects
// (union-find over object identities), in first-index order
std::vector<std::vector<int>>
groups_of (int count, const std::vector<std::vector<const void *>> &objs)
{
  std::vector<int> parent (count);
  std::iota (parent.begin (), parent.end (), 0);
  auto find = [&] (int i) {
    while (parent[i] != i) i = parent[i] = parent[parent[i]];
    return i;
  };
  std::map<const void *, int> owner;
  for (int i = 0; i < count; i++)
    for (const void *o : objs[i]) {
      if (!o) continue;
      auto it = owner.find (o);
      if (it == owner.end ()) owner[o] = i;
      else parent[find (i)] = find (it->second);
    }
  std::map<int, std::vector<int>> by_root;
  for (int i = 0; i < count; i++) by_root[find (i)].push_back (i);
  std::vector<std::vector<int>> out;
  for (auto &kv : by_root) out.push_back (kv.second);
  std::sort (out.begin (), out.end (),
             [] (const std::vector<int> &a, const std::vector<int> &b) {
               return a[0] < b[0];
             });
  return out;
}

// run job(i) for every instance on `workers` threads, group by group; the
// first failure (lowest instance index) is returned with its message
struct Failure { int index = -1; int code = 0; std::string msg; };

Failure
run_groups (const std::vector<std::vector<int>> &groups, int workers,
            const std::function<int (int)> &job)
{
  std::atomic<int> next (0);
  std::mutex mu;
  Failure first;
  auto work = [&] () {
    set_aux_overlap (false);             // the other workers fill the device
    // a block's partial last panel rides with the whole ones: the batch is
    // bound by launches, not bytes (256 uneven 2D instances: 2.59 -> 2.42 s)
    set_tail_split (false);
    for (;;) {
      const int gi = next.fetch_add (1);
      if (gi >= (int) groups.size ()) break;
      for (int i : groups[gi]) {
        const int rc = job (i);
        if (rc) {
          std::lock_guard<std::mutex> lk (mu);
          if (first.index < 0 || i < first.index) {
            first.index = i;
            first.code = rc;
            const char *m = lgh_petsc_last_error ();
            first.msg = (m && *m) ? m : lgh::last_error ();
          }
          break;                       // the rest of this group is skipped
        }
      }
    }
    dev_thread_exit ();
  };
  const int nt = std::max (1, std::min (workers, (int) groups.size ()));
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++) th.emplace_back (work);
  for (auto &t : th) t.join ();
  return first;
}

int
default_workers (int workers)
{
  return workers > 0 ? workers : 8;
}

}  // namespace

extern "C" int
lgh_prior_create_mat_batch (int count, Mat *Z, Vec *mass,
                            const lgh_prior_mat_opts_t *opts, int workers,
                            lgh_prior_t **priors)
{
  PetscCheck (count >= 0 && (count == 0 || (Z && mass && priors)), 0,
              PETSC_ERR_ARG_NULL, "lgh_prior_create_mat_batch: NULL argument");
  if (count == 0) return 0;
  if (dev_init ()) return lgh_petsc_error (PETSC_ERR_LIB, "%s", lgh::last_error ());
  std::vector<std::vector<const void *>> objs (count);
  for (int i = 0; i < count; i++) {
    priors[i] = nullptr;
    objs[i] = { Z[i], mass[i] };
  }
  const Failure f = run_groups (groups_of (count, objs), default_workers (workers),
                                [&] (int i) {
    return (int) lgh_prior_create_mat (Z[i], mass[i], opts ? &opts[i] : nullptr,
                                       &priors[i]);
  });
  if (f.index >= 0) {
    for (int i = 0; i < count; i++) { lgh_prior_destroy (priors[i]); priors[i] = nullptr; }
    return lgh_petsc_error (f.code, "lgh_prior_create_mat_batch: instance %d: %s",
                            f.index, f.msg.c_str ());
  }
  return 0;
}

extern "C" int
lgh_glr_compute_batch (int count, Mat *B, lgh_prior_t **priors,
                       const lgh_glr_opts_t *opts, const lgh_glr_xopts_t *xopts,
                       int workers, lgh_glr_t **glrs, lgh_glr_report_t *reports)
{
  PetscCheck (count >= 0 && (count == 0 || (B && priors && glrs)), 0,
              PETSC_ERR_ARG_NULL, "lgh_glr_compute_batch: NULL argument");
  if (count == 0) return 0;
  if (dev_init ()) return lgh_petsc_error (PETSC_ERR_LIB, "%s", lgh::last_error ());
  std::vector<std::vector<const void *>> objs (count);
  for (int i = 0; i < count; i++) {
    glrs[i] = nullptr;
    objs[i] = { B[i], priors[i] };
  }
  const Failure f = run_groups (groups_of (count, objs), default_workers (workers),
                                [&] (int i) {
    return (int) lgh_glr_compute_x (B[i], priors[i], opts ? &opts[i] : nullptr,
                                    xopts, &glrs[i],
                                    reports ? &reports[i] : nullptr);
  });
  if (f.index >= 0) {
    for (int i = 0; i < count; i++) { lgh_glr_destroy (glrs[i]); glrs[i] = nullptr; }
    return lgh_petsc_error (f.code, "lgh_glr_compute_batch: instance %d: %s",
                            f.index, f.msg.c_str ());
  }
  return 0;
}
