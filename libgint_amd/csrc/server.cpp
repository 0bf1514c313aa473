// Device-resident request loop for K-builds of many clients (the MPI ranks
// or OpenMP threads of independent SCF runs that would otherwise each hold
// their own copy of the library).  Nothing of this is in the reference,
// whose K-build is one blocking call sequence per caller (reference
// src/libGint.cpp:436-653); it serves the same per-client call protocol.
//
//  * A client is a handle whose screened call stream was recorded as a
//    resident plan (libgint_amd_plan_begin / plan_end).  The server gives it
//    a stream of its own and keeps its density P and exchange matrix K on
//    the device between requests.
//  * Requests (P in, K out, host arrays) queue in arrival order.  A worker
//    thread takes the oldest request whose client is idle and fits the
//    device budget (the loop is written for several workers whose requests
//    overlap on the GPU; one runs, see the constructor).
//  * Admission: the device bytes of the resident clients (plan records,
//    work buffers, tables, P and K) stay within the budget; an idle client is
//    evicted (plan copied to the host, device memory freed) when a request
//    of another client needs the room, least recently used first.
//  * Latency of a request: arrival (submit) to K copied to the host.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "engine.hpp"

namespace gint {

#define SCHECK(x)                                                     \
   do {                                                               \
      hipError_t e_ = (x);                                            \
      if (e_ != hipSuccess) fatal(#x ": " __FILE__, (int)e_);         \
   } while (0)

void launch_zero_f64(double *p, size_t n, void *stream);

namespace {

using Clock = std::chrono::steady_clock;

struct Client {
   Engine *e = nullptr;
   uint64_t n = 0;          // P / K doubles
   double *dP = nullptr, *dK = nullptr;
   bool resident = false;   // plan, buffers, P and K on the device
   bool busy = false;       // a request (or its eviction) is in progress
   uint64_t bytes = 0;      // device bytes while resident (measured; an estimate until then)
   uint64_t last_use = 0;
};

struct Request {
   int client;
   const double *P;
   double *K;
   Clock::time_point arrival, start, done;
   int state = 0;   // 0 queued, 1 running, 2 done, -1 refused (larger than the budget)
};

}  // namespace

class Server {
 public:
   // One worker thread, whatever `slots` asks: with four workers (four
   // clients' K-builds in flight at once) a 400-request run stopped twice
   // on the GPU after 111 and 151 requests -- four requests running, the
   // loop's mutex free, the workers blocked in the runtime or on the GPU --
   // and the cause is not found; concurrent workers stay off until it is.
   Server(size_t budget, int slots) : budget_(budget)
   {
      (void)slots;
      workers_.emplace_back([this] { work(); });
   }
   ~Server()
   {
      {
         std::unique_lock<std::mutex> lk(mu_);
         stop_ = true;
      }
      cv_.notify_all();
      for (auto &t : workers_) t.join();
      for (auto &c : clients_) {
         if (c.dP) (void)hipFree(c.dP);
         if (c.dK) (void)hipFree(c.dK);
      }
   }

   int add_client(Engine *e, uint64_t n)
   {
      if (e->plan_max_index() >= n) fatal("server_add_client: the plan's offsets reach beyond P_size", (long long)n);
      std::unique_lock<std::mutex> lk(mu_);
      e->use_private_stream();
      e->set_graph(false);   // no capture while other workers allocate
      if (e->plan_evicted()) e->plan_restore();   // measured while resident
      Client c;
      c.e = e;
      c.n = n;
      // the plan is on the device since plan_end: count it now
      c.resident = false;
      c.bytes = e->device_bytes() + 2 * 8 * n + e->work_estimate();
      clients_.push_back(c);
      const int id = (int)clients_.size() - 1;
      // it arrives resident (plan_end uploaded it): evict it when it does
      // not fit next to the clients already resident
      lk.unlock();
      e->plan_evict();
      return id;
   }

   uint64_t client_bytes(int client)
   {
      std::unique_lock<std::mutex> lk(mu_);
      return (client >= 0 && client < (int)clients_.size()) ? clients_[client].bytes : 0;
   }

   void set_budget(size_t b)
   {
      {
         std::unique_lock<std::mutex> lk(mu_);
         budget_ = b;
      }
      cv_.notify_all();
   }

   int64_t submit(int client, const double *P, double *K)
   {
      std::unique_lock<std::mutex> lk(mu_);
      if (client < 0 || client >= (int)clients_.size()) fatal("server_submit: unknown client", client);
      Request r;
      r.client = client;
      r.P = P;
      r.K = K;
      r.arrival = Clock::now();
      reqs_.push_back(r);
      queue_.push_back((int64_t)reqs_.size() - 1);
      lk.unlock();
      cv_.notify_all();
      return (int64_t)reqs_.size() - 1;
   }

   int wait(int64_t id, double out[3])
   {
      std::unique_lock<std::mutex> lk(mu_);
      if (id < 0 || id >= (int64_t)reqs_.size()) return -2;
      done_cv_.wait(lk, [&] { return reqs_[id].state == 2 || reqs_[id].state < 0; });
      const Request &r = reqs_[id];
      if (r.state < 0) return -1;
      out[0] = std::chrono::duration<double>(r.done - r.arrival).count();
      out[1] = std::chrono::duration<double>(r.start - r.arrival).count();
      out[2] = std::chrono::duration<double>(r.done - r.start).count();
      return 0;
   }

   void state(int64_t out[8])
   {
      std::unique_lock<std::mutex> lk(mu_);
      int64_t run = 0, done = 0, refused = 0, res = 0, busy = 0;
      for (const Request &r : reqs_) {
         run += r.state == 1;
         done += r.state == 2;
         refused += r.state < 0;
      }
      for (const Client &c : clients_) {
         res += c.resident;
         busy += c.busy;
      }
      out[0] = (int64_t)queue_.size();
      out[1] = run;
      out[2] = done;
      out[3] = refused;
      out[4] = res;
      out[5] = busy;
      out[6] = (int64_t)used_;
      out[7] = (int64_t)budget_;
   }

   void stats(double out[8])
   {
      std::unique_lock<std::mutex> lk(mu_);
      std::vector<double> lat;
      Clock::time_point t0 = Clock::time_point::max(), t1 = Clock::time_point::min();
      for (const Request &r : reqs_) {
         if (r.state != 2) continue;
         lat.push_back(std::chrono::duration<double>(r.done - r.arrival).count());
         t0 = std::min(t0, r.arrival);
         t1 = std::max(t1, r.done);
      }
      std::sort(lat.begin(), lat.end());
      auto pct = [&](double p) {
         if (lat.empty()) return 0.0;
         const size_t i = (size_t)std::min<double>((double)lat.size() - 1, std::ceil(p * (double)lat.size()) - 1);
         return lat[i];
      };
      double mean = 0;
      for (double x : lat) mean += x;
      out[0] = (double)lat.size();
      out[1] = pct(0.50);
      out[2] = pct(0.99);
      out[3] = lat.empty() ? 0.0 : mean / (double)lat.size();
      out[4] = lat.empty() ? 0.0 : std::chrono::duration<double>(t1 - t0).count();
      out[5] = (double)peak_used_;
      out[6] = (double)evictions_;
      out[7] = (double)admissions_;
   }

 private:
   // the oldest queued request whose client is idle and fits (evicting idle
   // clients, least recently used first); -1 if none.  Under mu_.
   int64_t pick(std::vector<int> &evict)
   {
      for (size_t qi = 0; qi < queue_.size(); qi++) {
         const int64_t id = queue_[qi];
         Client &c = clients_[reqs_[id].client];
         if (c.busy) continue;
         if (!c.resident) {
            if (c.bytes > budget_) {
               reqs_[id].state = -1;   // can never fit
               queue_.erase(queue_.begin() + (long)qi);
               done_cv_.notify_all();
               return pick(evict);
            }
            uint64_t free = used_ >= budget_ ? 0 : budget_ - used_;
            std::vector<int> idle;
            for (int i = 0; i < (int)clients_.size(); i++)
               if (clients_[i].resident && !clients_[i].busy) idle.push_back(i);
            std::sort(idle.begin(), idle.end(), [&](int a, int b) { return clients_[a].last_use < clients_[b].last_use; });
            std::vector<int> ev;
            for (int i : idle) {
               if (free >= c.bytes) break;
               ev.push_back(i);
               free += clients_[i].bytes;
            }
            if (free < c.bytes) continue;   // wait for running clients
            for (int i : ev) {
               clients_[i].resident = false;
               clients_[i].busy = true;     // being evicted
               used_ -= clients_[i].bytes;
               evictions_++;
            }
            evict = ev;
            c.resident = true;
            used_ += c.bytes;
            peak_used_ = std::max(peak_used_, used_);
            admissions_++;
         }
         c.busy = true;
         c.last_use = ++tick_;
         queue_.erase(queue_.begin() + (long)qi);
         return id;
      }
      return -1;
   }

   void work()
   {
      for (;;) {
         int64_t id;
         std::vector<int> evict;
         // a copy of the request and pointers to the clients, taken under
         // mu_: indexing a deque while submit / add_client push_back reads
         // its block map, which growing reallocates (a worker could take a
         // freed map, the wrong request or none); pointers to elements stay
         // valid
         Request req;
         std::vector<Client *> ev;
         Client *cp = nullptr;
         {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] {
               if (stop_) return true;
               id = pick(evict);
               return id >= 0;
            });
            if (stop_ && id < 0) return;
            if (id < 0) continue;
            reqs_[id].state = 1;
            reqs_[id].start = Clock::now();
            req = reqs_[id];
            cp = &clients_[req.client];
            for (int i : evict) ev.push_back(&clients_[i]);
         }
         for (Client *xp : ev) {
            Client &x = *xp;
            x.e->plan_evict();
            if (x.dP) SCHECK(hipFree(x.dP));
            if (x.dK) SCHECK(hipFree(x.dK));
            x.dP = x.dK = nullptr;
            std::unique_lock<std::mutex> lk(mu_);
            x.busy = false;
         }
         const Request &r = req;
         Client &c = *cp;
         if (c.e->plan_evicted()) c.e->plan_restore();
         if (!c.dP) {
            SCHECK(hipMalloc((void **)&c.dP, sizeof(double) * c.n));
            SCHECK(hipMalloc((void **)&c.dK, sizeof(double) * c.n));
         }
         hipStream_t q = (hipStream_t)c.e->stream();
         SCHECK(hipMemcpyAsync(c.dP, r.P, sizeof(double) * c.n, hipMemcpyHostToDevice, q));
         launch_zero_f64(c.dK, c.n, q);
         c.e->plan_execute_into(c.dP, nullptr, c.dK, nullptr, 1, c.n);
         SCHECK(hipMemcpyAsync(r.K, c.dK, sizeof(double) * c.n, hipMemcpyDeviceToHost, q));
         c.e->synchronize();
         const uint64_t measured = c.e->device_bytes() + 2 * 8 * c.n;
         {
            std::unique_lock<std::mutex> lk(mu_);
            Request &rd = reqs_[id];
            rd.done = Clock::now();
            rd.state = 2;
            // the footprint the budget counts from now on
            used_ = used_ - c.bytes + measured;
            c.bytes = measured;
            peak_used_ = std::max(peak_used_, used_);
            c.busy = false;
         }
         done_cv_.notify_all();
         cv_.notify_all();
      }
   }

   size_t budget_;
   uint64_t used_ = 0, peak_used_ = 0, tick_ = 0, evictions_ = 0, admissions_ = 0;
   bool stop_ = false;
   std::mutex mu_;
   std::condition_variable cv_, done_cv_;
   std::deque<Client> clients_;   // deques: references stay valid while others are added
   std::deque<Request> reqs_;
   std::deque<int64_t> queue_;
   std::vector<std::thread> workers_;
};

}  // namespace gint

extern "C" {

void *libgint_amd_server_create(int max_mem_MiB, int slots)
{
   return new gint::Server((size_t)std::max(max_mem_MiB, 1) << 20, slots);
}
int libgint_amd_server_add_client(void *srv, void *handle, int P_size)
{
   return static_cast<gint::Server *>(srv)->add_client(static_cast<gint::Engine *>(handle), (uint64_t)P_size);
}
int64_t libgint_amd_server_submit(void *srv, int client, const double *P, double *K)
{
   return static_cast<gint::Server *>(srv)->submit(client, P, K);
}
int libgint_amd_server_wait(void *srv, int64_t request, double out[3])
{
   return static_cast<gint::Server *>(srv)->wait(request, out);
}
void libgint_amd_server_stats(void *srv, double out[8]) { static_cast<gint::Server *>(srv)->stats(out); }
int64_t libgint_amd_server_client_bytes(void *srv, int client)
{
   return (int64_t) static_cast<gint::Server *>(srv)->client_bytes(client);
}
void libgint_amd_server_set_budget(void *srv, int max_mem_MiB)
{
   static_cast<gint::Server *>(srv)->set_budget((size_t)std::max(max_mem_MiB, 1) << 20);
}
void libgint_amd_server_state(void *srv, int64_t out[8]) { static_cast<gint::Server *>(srv)->state(out); }
void libgint_amd_server_destroy(void *srv) { delete static_cast<gint::Server *>(srv); }

}  // extern "C"
