// Serve front-end: the reference's TCP protocol in front of the search,
// with every search run as fixed-size batches so a captured search
// (hipGraph) is replayed for each of them.
//
// Protocol (little-endian), src/master_server.h:5-8, src/master_server.cpp:
// 105-197, src/README.md:24-44:
//   INSERT phase (first connection, when enabled), src/query.cpp:555-660:
//     [u8 op=1][u64 n][n*d f32] -> [u8 status]    rows buffered
//     [u8 op=2]                 -> [u8 status]    build: train (when the index
//                                                 has no centroids) + add
//     [u8 op=0]                                   abort the INSERT phase
//   QUERY phase (every later connection, one client at a time):
//     [u64 nq][u64 k][nq*d f32] -> [nq*k f32 distances][nq*k i64 labels]
//     [u64 0][u64 0]                              shutdown
// Status bytes: 0 OK, 1 error.  A query with k outside [1, 1024], or larger
// than the configured request limit, closes that connection.
//
// The reference hands each request to an MPI master loop through a job
// queue (src/query.cpp:702-799); here the listener thread runs the batches
// itself: a request's rows are cut into batches of `batch` rows (the last
// one padded with earlier rows), each copied into pinned memory, uploaded,
// searched on the device by the captured search, and its first rows'
// results copied back.
#include <arpa/inet.h>
#include <netinet/in.h>
#include <netinet/tcp.h>
#include <sys/socket.h>
#include <unistd.h>

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "index_impl.hpp"

namespace hivf {
namespace {

constexpr uint8_t kOpShutdown = 0x00, kOpInsert = 0x01, kOpBuildDone = 0x02;
constexpr uint8_t kStatusOk = 0, kStatusError = 1;

bool send_all(int fd, const void* buf, size_t n) {
    const char* p = static_cast<const char*>(buf);
    while (n > 0) {
        ssize_t s = ::send(fd, p, n, MSG_NOSIGNAL);
        if (s < 0 && errno == EINTR) continue;
        if (s <= 0) return false;
        p += s;
        n -= (size_t)s;
    }
    return true;
}

bool recv_all(int fd, void* buf, size_t n) {
    char* p = static_cast<char*>(buf);
    while (n > 0) {
        ssize_t s = ::recv(fd, p, n, 0);
        if (s < 0 && errno == EINTR) continue;
        if (s <= 0) return false;
        p += s;
        n -= (size_t)s;
    }
    return true;
}

// What the server runs: a batch of rows, and the build at BUILD_DONE.
struct Runner {
    virtual ~Runner() = default;
    // rows <= batch; x holds rows * d floats; D / I receive rows * k results
    virtual int search(int64_t rows, const float* x, int64_t k, float* D, int64_t* I) = 0;
    virtual int build(int64_t n, const float* x) = 0;
};

// The library's own search on one index: pinned staging, device buffers of
// one batch, the captured search replayed per batch.
struct GpuRunner : Runner {
    hivf_index* h;
    hivf_search_params p;
    int32_t train_niter;
    int64_t d, batch;
    float* hx = nullptr;     // pinned [batch x d]
    float* hD = nullptr;     // pinned [batch x kcap]
    int64_t* hI = nullptr;
    float* dx = nullptr;     // device [batch x d]
    float* dD = nullptr;     // device [batch x kcap]
    int64_t* dI = nullptr;
    int64_t kcap = 0;

    ~GpuRunner() override {
        (void)hipSetDevice(h->device);
        if (h->stream) (void)hipStreamSynchronize(h->stream);
        if (hx) (void)hipHostFree(hx);
        if (hD) (void)hipHostFree(hD);
        if (hI) (void)hipHostFree(hI);
        dev_free(dx);
        dev_free(dD);
        dev_free(dI);
    }
    int init() {
        HIP_TRY(h, hipSetDevice(h->device));
        HIP_TRY(h, hipHostMalloc(reinterpret_cast<void**>(&hx), (size_t)batch * d * sizeof(float)));
        std::memset(hx, 0, (size_t)batch * d * sizeof(float));
        HIP_TRY(h, hipMalloc(reinterpret_cast<void**>(&dx), (size_t)batch * d * sizeof(float)));
        return HIVF_OK;
    }
    int ensure_k(int64_t k) {
        if (k <= kcap) return HIVF_OK;
        HIP_TRY(h, hipStreamSynchronize(h->stream));
        if (hD) (void)hipHostFree(hD);
        if (hI) (void)hipHostFree(hI);
        hD = nullptr;
        hI = nullptr;
        dev_free(dD);
        dev_free(dI);
        dD = nullptr;
        dI = nullptr;
        kcap = 0;
        HIP_TRY(h, hipHostMalloc(reinterpret_cast<void**>(&hD), (size_t)batch * k * sizeof(float)));
        HIP_TRY(h, hipHostMalloc(reinterpret_cast<void**>(&hI), (size_t)batch * k * sizeof(int64_t)));
        HIP_TRY(h, hipMalloc(reinterpret_cast<void**>(&dD), (size_t)batch * k * sizeof(float)));
        HIP_TRY(h, hipMalloc(reinterpret_cast<void**>(&dI), (size_t)batch * k * sizeof(int64_t)));
        kcap = k;
        return HIVF_OK;
    }
    int search(int64_t rows, const float* x, int64_t k, float* D, int64_t* I) override {
        HIP_TRY(h, hipSetDevice(h->device));
        int rc = ensure_k(k);
        if (rc) return rc;
        std::memcpy(hx, x, (size_t)rows * d * sizeof(float));
        HIP_TRY(h, hipMemcpyAsync(dx, hx, (size_t)batch * d * sizeof(float), hipMemcpyHostToDevice, h->stream));
        rc = hivf_index_search_device(h, batch, dx, k, dD, dI, &p, nullptr, h->stream);
        if (rc) return rc;
        HIP_TRY(h, hipMemcpyAsync(hD, dD, (size_t)rows * k * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(h, hipMemcpyAsync(hI, dI, (size_t)rows * k * sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(h, hipStreamSynchronize(h->stream));
        std::memcpy(D, hD, (size_t)rows * k * sizeof(float));
        std::memcpy(I, hI, (size_t)rows * k * sizeof(int64_t));
        return HIVF_OK;
    }
    int build(int64_t n, const float* x) override {
        if (!h->have_centroids) {
            int rc = hivf_index_train_kmeans(h, n, x, train_niter, 6666, 256);
            if (rc) return rc;
        }
        return hivf_index_add(h, n, x);
    }
};

// A caller-supplied search (for example a list-sharded one) behind the same
// transport and batching.
struct FnRunner : Runner {
    hivf_search_fn fn;
    hivf_build_fn bfn;
    void* user;
    int search(int64_t rows, const float* x, int64_t k, float* D, int64_t* I) override {
        return fn(user, rows, x, k, D, I);
    }
    int build(int64_t n, const float* x) override { return bfn ? bfn(user, n, x) : HIVF_EUNSUPPORTED; }
};

}  // namespace
}  // namespace hivf

struct hivf_server {
    std::unique_ptr<hivf::Runner> runner;
    int64_t d = 0, batch = 0, max_request_rows = 0;
    bool insert_phase = false;
    int listen_fd = -1;
    int port = 0;
    std::atomic<int> client_fd{-1};
    std::atomic<bool> stop{false};
    std::thread th;
    std::mutex mu;
    int status = HIVF_OK;
    std::string err;
    hivf_server_stats st{};

    void set_error(int code, const std::string& m) {
        std::lock_guard<std::mutex> lk(mu);
        status = code;
        err = m;
    }

    // INSERT phase on one connection; true once BUILD_DONE succeeded.
    bool insert_loop(int fd) {
        std::vector<float> rows;
        int64_t total = 0;
        for (;;) {
            uint8_t op;
            if (!recv_all_(fd, &op, 1)) return false;
            if (op == hivf::kOpShutdown) return false;
            if (op == hivf::kOpInsert) {
                uint64_t n = 0;
                if (!recv_all_(fd, &n, sizeof(n))) return false;
                if (n > (uint64_t)max_request_rows || total + (int64_t)n > INT32_MAX) {
                    send_status(fd, hivf::kStatusError);
                    return false;
                }
                rows.resize((size_t)(total + (int64_t)n) * d);
                if (!recv_all_(fd, rows.data() + (size_t)total * d, (size_t)n * d * sizeof(float))) return false;
                total += (int64_t)n;
                if (!send_status(fd, hivf::kStatusOk)) return false;
            } else if (op == hivf::kOpBuildDone) {
                int rc = runner->build(total, rows.data());
                if (rc) set_error(rc, "build at BUILD_DONE failed");
                {
                    std::lock_guard<std::mutex> lk(mu);
                    st.rows_inserted += total;
                }
                if (!send_status(fd, rc ? hivf::kStatusError : hivf::kStatusOk)) return false;
                return rc == HIVF_OK;
            } else {
                if (!send_status(fd, hivf::kStatusError)) return false;
            }
        }
    }

    // QUERY phase on one connection; true when the client asked for shutdown.
    bool query_loop(int fd) {
        std::vector<float> x, D;
        std::vector<int64_t> I;
        for (;;) {
            uint64_t hdr[2];
            if (!recv_all_(fd, hdr, sizeof(hdr))) return false;
            const uint64_t nq = hdr[0], k = hdr[1];
            if (nq == 0) return true;  // src/master_server.cpp:148-152
            if (k < 1 || k > 1024 || nq > (uint64_t)max_request_rows) {
                set_error(HIVF_EINVAL, "request with k outside [1, 1024] or too many rows: connection closed");
                return false;
            }
            x.resize((size_t)nq * d);
            if (!recv_all_(fd, x.data(), x.size() * sizeof(float))) return false;
            D.resize((size_t)nq * k);
            I.resize((size_t)nq * k);
            int64_t batches = 0;
            for (uint64_t r0 = 0; r0 < nq; r0 += (uint64_t)batch) {
                const int64_t rows = (int64_t)std::min<uint64_t>((uint64_t)batch, nq - r0);
                int rc = runner->search(rows, x.data() + r0 * d, (int64_t)k, D.data() + r0 * k, I.data() + r0 * k);
                if (rc) {
                    set_error(rc, "search failed: connection closed");
                    return false;
                }
                ++batches;
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                st.requests += 1;
                st.queries += (int64_t)nq;
                st.batches += batches;
            }
            if (!send_all_(fd, D.data(), D.size() * sizeof(float))) return false;
            if (!send_all_(fd, I.data(), I.size() * sizeof(int64_t))) return false;
        }
    }

    void serve() {
        bool inserting = insert_phase;
        while (!stop.load()) {
            sockaddr_in ca{};
            socklen_t cl = sizeof(ca);
            int fd = ::accept(listen_fd, reinterpret_cast<sockaddr*>(&ca), &cl);
            if (fd < 0) {
                if (errno == EINTR) continue;
                break;  // closed by hivf_server_stop, or a listener error
            }
            int one = 1;
            setsockopt(fd, IPPROTO_TCP, TCP_NODELAY, &one, sizeof(one));
            client_fd.store(fd);
            bool shutdown_req = false;
            if (inserting) {
                if (insert_loop(fd)) inserting = false;
            } else {
                shutdown_req = query_loop(fd);
            }
            client_fd.store(-1);
            ::close(fd);
            {
                std::lock_guard<std::mutex> lk(mu);
                st.connections += 1;
            }
            if (shutdown_req) break;
        }
        stop.store(true);
    }

    bool recv_all_(int fd, void* b, size_t n) { return hivf::recv_all(fd, b, n); }
    bool send_all_(int fd, const void* b, size_t n) { return hivf::send_all(fd, b, n); }
    bool send_status(int fd, uint8_t s) { return hivf::send_all(fd, &s, 1); }
};

namespace {

hivf_server* start_server(std::unique_ptr<hivf::Runner> runner, int64_t d, const char* address, int port,
                          int64_t batch, int insert_phase, int64_t max_request_rows, int* status) {
    auto failed = [&](int code, const std::string& m) -> hivf_server* {
        hivf::set_global_error(m);
        if (status) *status = code;
        return nullptr;
    };
    auto s = std::make_unique<hivf_server>();
    s->runner = std::move(runner);
    s->d = d;
    s->batch = batch;
    s->insert_phase = insert_phase != 0;
    s->max_request_rows = max_request_rows > 0 ? max_request_rows : (int64_t(1) << 24);
    s->listen_fd = ::socket(AF_INET, SOCK_STREAM, 0);
    if (s->listen_fd < 0) return failed(HIVF_EIO, "hivf_server_start: socket() failed");
    int one = 1;
    setsockopt(s->listen_fd, SOL_SOCKET, SO_REUSEADDR, &one, sizeof(one));
    sockaddr_in addr{};
    addr.sin_family = AF_INET;
    addr.sin_port = htons((uint16_t)port);
    if (::inet_pton(AF_INET, address ? address : "0.0.0.0", &addr.sin_addr) != 1) {
        ::close(s->listen_fd);
        return failed(HIVF_EINVAL, "hivf_server_start: bad IPv4 address");
    }
    if (::bind(s->listen_fd, reinterpret_cast<sockaddr*>(&addr), sizeof(addr)) < 0 || ::listen(s->listen_fd, 1) < 0) {
        ::close(s->listen_fd);
        return failed(HIVF_EIO, "hivf_server_start: bind() / listen() failed on port " + std::to_string(port));
    }
    socklen_t al = sizeof(addr);
    ::getsockname(s->listen_fd, reinterpret_cast<sockaddr*>(&addr), &al);
    s->port = ntohs(addr.sin_port);
    hivf_server* raw = s.get();
    raw->th = std::thread([raw]() { raw->serve(); });
    if (status) *status = HIVF_OK;
    return s.release();
}

}  // namespace

extern "C" {

void hivf_serve_params_default(hivf_serve_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->batch = 4096;
    p->train_niter = 20;
    hivf_search_params_default(&p->search);
    p->search.algo = HIVF_ALGO_BATCHED;
}

hivf_server* hivf_server_start(hivf_index* h, const char* address, int port, const hivf_serve_params* params,
                               int* status) {
    HIVF_API_BEGIN
    if (!h || port < 0 || port > 65535) {
        hivf::set_global_error("hivf_server_start: null index or bad port");
        if (status) *status = HIVF_EINVAL;
        return nullptr;
    }
    hivf_serve_params p;
    hivf_serve_params_default(&p);
    if (params) p = *params;
    auto r = std::make_unique<hivf::GpuRunner>();
    r->h = h;
    r->p = p.search;
    r->p.use_graph = 1;
    r->p.nprobe_per_query = nullptr;
    r->train_niter = p.train_niter > 0 ? p.train_niter : 20;  // src/Index.cpp:389
    r->d = h->d;
    r->batch = p.batch > 0 ? p.batch : 4096;
    int rc = r->init();
    if (rc) {
        hivf::set_global_error("hivf_server_start: " + h->err);
        if (status) *status = rc;
        return nullptr;
    }
    const int64_t batch = r->batch;
    return start_server(std::move(r), h->d, address, port, batch, p.insert_phase, p.max_request_rows, status);
    HIVF_API_END_PTR(status)
}

hivf_server* hivf_server_start_fn(int64_t d, const char* address, int port, int64_t batch, int insert_phase,
                                  hivf_search_fn search, hivf_build_fn build, void* user, int* status) {
    HIVF_API_BEGIN
    if (d < 1 || !search || port < 0 || port > 65535) {
        hivf::set_global_error("hivf_server_start_fn: bad d, port or null search function");
        if (status) *status = HIVF_EINVAL;
        return nullptr;
    }
    auto r = std::make_unique<hivf::FnRunner>();
    r->fn = search;
    r->bfn = build;
    r->user = user;
    return start_server(std::move(r), d, address, port, batch > 0 ? batch : 4096, insert_phase, 0, status);
    HIVF_API_END_PTR(status)
}

int hivf_server_port(const hivf_server* s) { return s ? s->port : -1; }

int hivf_server_wait(hivf_server* s) {
    HIVF_API_BEGIN
    if (!s) return HIVF_EINVAL;
    if (s->th.joinable()) s->th.join();
    std::lock_guard<std::mutex> lk(s->mu);
    return s->status;
    HIVF_API_END(nullptr)
}

int hivf_server_stats_get(hivf_server* s, hivf_server_stats* out) {
    HIVF_API_BEGIN
    if (!s || !out) return HIVF_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    return HIVF_OK;
    HIVF_API_END(nullptr)
}

const char* hivf_server_error(hivf_server* s) {
    if (!s) return "";
    std::lock_guard<std::mutex> lk(s->mu);
    return s->err.c_str();
}

void hivf_server_stop(hivf_server* s) {
    HIVF_API_BEGIN
    if (!s) return;
    s->stop.store(true);
    ::shutdown(s->listen_fd, SHUT_RDWR);
    const int c = s->client_fd.load();
    if (c >= 0) ::shutdown(c, SHUT_RDWR);
    if (s->th.joinable()) s->th.join();
    ::close(s->listen_fd);
    delete s;
    HIVF_API_END_VOID
}

}  // extern "C"
