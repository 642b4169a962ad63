// sq_serve.hip - a serving loop of SCF solves on one device.
//
// Requests (a molecule, a run of geometry steps, charge and multiplicity,
// an arrival time) are admitted while a slot is free and the summed device
// footprint of the requests in flight stays within the memory budget; a
// request that does not fit waits for the next one to finish.  Each slot is
// a host thread with its own device-resident SCF context (sqscf_*) and so
// its own stream: the J/K builds, Fock updates and density steps of the
// requests in flight run on the device at the same time, and a request
// arriving while others run starts as soon as it is admitted, not when the
// others end.  The steps of one request follow the reference's geometry
// loop (src/qmd.c:564-667, src/optimize.c:283-340): the first step starts
// from the DIAG guess, later ones from the previous step's orbitals
// S-normalised in the new geometry (SCFGUESS_CACB, src/qmd.c:585-589), and
// the J/K context of the slot moves to the new geometry instead of being
// built again (sqjk_update_geometry).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sq_jk.h"
#include "../../include/sq_scf.h"
#include "../../include/sq_serve.h"

namespace {

thread_local std::string g_srv_err;
int srv_fail(int code, const std::string& m) {
  g_srv_err = m;
  return code;
}

using Clock = std::chrono::steady_clock;

struct Request {
  int natom = 0, nstep = 0, charge = 0, mult = 1;
  std::vector<int> Z;
  std::vector<double> xyz;  // nstep x natom x 3, bohr
  double arrival = 0.0;
  // results
  std::vector<double> energy;
  std::vector<int> iterations;
  double start = -1.0, finish = -1.0, footprint = 0.0;
  int nBasis = 0, converged = 0, iters_total = 0, error = 0;
  std::string errmsg;
};

}  // namespace

struct sqsrv {
  sqsrv_config cfg{};
  int dbSize = 0;
  const GTOBasisSet_t* db = nullptr;
  std::vector<Request> req;
  // scheduler state
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<int> queue;  // admitted, waiting for a slot thread
  int in_flight = 0, finished = 0;
  double mem_in_flight = 0.0, mem_peak = 0.0;
  int peak_in_flight = 0;
  bool stop = false;
  Clock::time_point t0;
};

namespace {

double since(const sqsrv* s) { return std::chrono::duration<double>(Clock::now() - s->t0).count(); }

void free_basis(GTOBasis_t* g, int n) {
  if (!g) return;
  for (int i = 0; i < n; ++i) {
    std::free(g[i].exp);
    std::free(g[i].coef);
    std::free(g[i].norm);
  }
  std::free(g);
}

// One request, all of its geometry steps, on the slot's SCF context.
void serve_one(sqsrv* s, Request& r, sqscf_ctx* scf) {
  const sqsrv_config& c = s->cfg;
  int ne = -r.charge;
  for (int z : r.Z) ne += z;
  const int nEA = (ne + r.mult - 1) / 2, nEB = (ne - r.mult + 1) / 2;
  if (nEA < 0 || nEB < 0 || nEA + nEB != ne) {
    r.error = SQJK_EARG;
    r.errmsg = "electron count and multiplicity do not match";
    return;
  }
  std::vector<double> x(r.natom), y(r.natom), z(r.natom);
  for (int k = 0; k < r.nstep; ++k) {
    for (int a = 0; a < r.natom; ++a) {
      const double* p = &r.xyz[((size_t)k * r.natom + a) * 3];
      x[a] = p[0]; y[a] = p[1]; z[a] = p[2];
    }
    Molecule_t mol{};
    mol.nAtom = r.natom;
    mol.Q = r.charge;
    mol.Z = r.Z.data();
    mol.x = x.data(); mol.y = y.data(); mol.z = z.data();
    int n = 0;
    GTOBasis_t* gto = genBasis(&mol, &n, s->dbSize, s->db);
    r.nBasis = n;
    const size_t nn = (size_t)n * n;
    std::vector<double> T(nn), V(nn), S(nn);
    int rc = sqjk_one_electron(n, gto, &mol, c.device, T.data(), V.data(), S.data());
    if (!rc) {
      for (size_t e = 0; e < nn; ++e) T[e] += V[e];  // H
      rc = sqscf_set_system(scf, n, gto, T.data(), S.data());
    }
    if (!rc) {
      if (k == 0) {
        // DIAG guess (src/uhf.c:1384-1399)
        double tr = 0.0;
        for (int i = 0; i < n; ++i) tr += S[(size_t)i * n + i];
        std::fill(T.begin(), T.end(), 0.0);
        std::fill(V.begin(), V.end(), 0.0);
        for (int i = 0; i < n; ++i) {
          T[(size_t)i * n + i] = nEA / tr;
          V[(size_t)i * n + i] = nEB / tr;
        }
        rc = sqscf_set_density(scf, T.data(), V.data());
      } else {
        rc = sqscf_guess_from_orbitals(scf, nEA, nEB);
      }
    }
    sqscf_result res{};
    if (!rc) {
      sqscf_params p{};
      p.restricted = r.mult == 1 ? 1 : 0;
      p.nEA = nEA;
      p.nEB = nEB;
      p.SCFConv = c.SCFConv;
      p.SCFMax = c.SCFMax;
      p.SCFDrag = c.SCFDrag;
      p.threeStep = c.threeStep;
      p.convMethod = CONVMETHOD_DIIS4;
      p.Econst = nuclei_coulomb(&mol);
      p.mixedThreshold = c.mixedThreshold;
      rc = sqscf_run(scf, &p, &res);
    }
    free_basis(gto, n);
    if (rc) {
      r.error = rc;
      r.errmsg = std::string("step ") + std::to_string(k) + " (" + std::to_string(n) + " functions): " +
                 (rc == SQJK_OK ? "" : (sqscf_errmsg()[0] ? sqscf_errmsg() : sqjk_errmsg()));
      return;
    }
    r.energy[k] = res.converged ? res.energy : 0.0;
    r.iterations[k] = res.iterations;
    r.iters_total += res.iterations;
    r.converged += res.converged;
  }
}

void slot_thread(sqsrv* s) {
  sqscf_ctx* scf = nullptr;
  if (hipSetDevice(s->cfg.device) != hipSuccess || sqscf_create(s->cfg.device, &scf)) scf = nullptr;
  for (;;) {
    int id;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
      if (s->queue.empty()) break;
      id = s->queue.front();
      s->queue.pop_front();
    }
    Request& r = s->req[id];
    r.start = since(s);
    if (scf) {
      serve_one(s, r, scf);
    } else {
      r.error = SQJK_EHIP;
      r.errmsg = "SCF context creation failed";
    }
    r.finish = since(s);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->in_flight--;
      s->finished++;
      s->mem_in_flight -= r.footprint;
    }
    s->cv_done.notify_all();
  }
  sqscf_destroy(scf);
}

// device footprint of a request (its first geometry; the steps keep the basis)
double request_footprint(const sqsrv* s, const Request& r) {
  Molecule_t mol{};
  std::vector<double> x(r.natom), y(r.natom), z(r.natom);
  for (int a = 0; a < r.natom; ++a) {
    x[a] = r.xyz[3 * a]; y[a] = r.xyz[3 * a + 1]; z[a] = r.xyz[3 * a + 2];
  }
  std::vector<int> Z = r.Z;
  mol.nAtom = r.natom;
  mol.Q = r.charge;
  mol.Z = Z.data();
  mol.x = x.data(); mol.y = y.data(); mol.z = z.data();
  int n = 0;
  GTOBasis_t* gto = genBasis(&mol, &n, s->dbSize, s->db);
  const long long b = sqjk_footprint_bytes(n, gto, r.mult == 1 ? 0 : 1, 1);
  free_basis(gto, n);
  return b < 0 ? 0.0 : (double)b;
}

}  // namespace

extern "C" {

const char* sqsrv_errmsg(void) { return g_srv_err.c_str(); }

int sqsrv_create(const sqsrv_config* cfg, int dbSize, const struct GTOBasisSet_t* db, sqsrv** out) {
  if (!cfg || !out || dbSize <= 0 || !db || cfg->slots < 1 || cfg->slots > 64 || !(cfg->memBudget > 0.0) ||
      cfg->SCFMax < 1 || !(cfg->SCFConv > 0.0))
    return srv_fail(SQJK_EARG, "sqsrv_create: invalid arguments");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return srv_fail(SQJK_ENODEV, "no HIP device is visible; this library has no CPU path");
  if (cfg->device < 0 || cfg->device >= ndev) return srv_fail(SQJK_EARG, "sqsrv_create: device out of range");
  sqsrv* s = new sqsrv();
  s->cfg = *cfg;
  s->dbSize = dbSize;
  s->db = db;
  *out = s;
  return SQJK_OK;
}

int sqsrv_submit(sqsrv* s, int nAtom, const int* Z, int nStep, const double* xyz, int charge, int multiplicity,
                 double arrival) {
  if (!s || nAtom <= 0 || !Z || nStep <= 0 || !xyz || multiplicity < 1 || arrival < 0.0)
    return srv_fail(SQJK_EARG, "sqsrv_submit: invalid arguments");
  Request r;
  r.natom = nAtom;
  r.nstep = nStep;
  r.charge = charge;
  r.mult = multiplicity;
  r.Z.assign(Z, Z + nAtom);
  r.xyz.assign(xyz, xyz + (size_t)nStep * nAtom * 3);
  r.arrival = arrival;
  r.energy.assign(nStep, 0.0);
  r.iterations.assign(nStep, 0);
  s->req.push_back(std::move(r));
  return (int)s->req.size() - 1;
}

int sqsrv_run(sqsrv* s) {
  if (!s) return srv_fail(SQJK_EARG, "sqsrv_run: null server");
  const int nreq = (int)s->req.size();
  std::vector<int> order(nreq);
  for (int i = 0; i < nreq; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return s->req[a].arrival < s->req[b].arrival; });
  for (Request& r : s->req) r.footprint = request_footprint(s, r);
  s->in_flight = s->finished = 0;
  s->mem_in_flight = s->mem_peak = 0.0;
  s->peak_in_flight = 0;
  s->stop = false;
  s->queue.clear();
  std::vector<std::thread> pool;
  s->t0 = Clock::now();
  for (int k = 0; k < s->cfg.slots; ++k) pool.emplace_back(slot_thread, s);
  size_t next = 0;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    while (s->finished < nreq) {
      // admit in arrival order while a slot and the memory budget allow; a
      // request larger than the whole budget runs alone
      while (next < order.size()) {
        Request& r = s->req[order[next]];
        if (r.arrival > since(s) || s->in_flight >= s->cfg.slots) break;
        if (s->in_flight > 0 && s->mem_in_flight + r.footprint > s->cfg.memBudget) break;
        s->queue.push_back(order[next]);
        s->in_flight++;
        s->mem_in_flight += r.footprint;
        s->mem_peak = std::max(s->mem_peak, s->mem_in_flight);
        s->peak_in_flight = std::max(s->peak_in_flight, s->in_flight);
        ++next;
        s->cv_work.notify_one();
      }
      if (s->finished >= nreq) break;
      // wait for a completion or the next arrival
      if (next < order.size() && s->in_flight < s->cfg.slots &&
          !(s->in_flight > 0 && s->mem_in_flight + s->req[order[next]].footprint > s->cfg.memBudget)) {
        const double dt = s->req[order[next]].arrival - since(s);
        if (dt > 0) s->cv_done.wait_for(lk, std::chrono::duration<double>(dt));
      } else {
        s->cv_done.wait(lk);
      }
    }
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (std::thread& t : pool) t.join();
  for (const Request& r : s->req)
    if (r.error) return srv_fail(r.error, "request failed: " + r.errmsg);
  return SQJK_OK;
}

int sqsrv_result(const sqsrv* s, int id, double* energies, int* iterations, sqsrv_stats* st) {
  if (!s || id < 0 || id >= (int)s->req.size()) return srv_fail(SQJK_EARG, "sqsrv_result: invalid request id");
  const Request& r = s->req[id];
  if (energies) std::copy(r.energy.begin(), r.energy.end(), energies);
  if (iterations) std::copy(r.iterations.begin(), r.iterations.end(), iterations);
  if (st) {
    st->nstep = r.nstep;
    st->converged_steps = r.converged;
    st->nBasis = r.nBasis;
    st->iterations = r.iters_total;
    st->arrival = r.arrival;
    st->start = r.start;
    st->finish = r.finish;
    st->footprint = r.footprint;
    st->peak_in_flight = s->peak_in_flight;
    st->peak_memory = s->mem_peak;
  }
  return SQJK_OK;
}

void sqsrv_destroy(sqsrv* s) { delete s; }

}  // extern "C"
