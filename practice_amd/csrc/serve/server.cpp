// Query server (serial worker).
#include "server.h"

#include <algorithm>
#include <sstream>

namespace db {

QueryServer::QueryServer(DatabaseInstance& db, const std::string& options) : db_(db) {
    (void)options;
    worker_ = std::thread([this] { loop(); });
}

QueryServer::~QueryServer() {
    {
        std::lock_guard<std::mutex> lock(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    if (worker_.joinable()) worker_.join();
}

long long QueryServer::submit(const std::string& sql) {
    std::lock_guard<std::mutex> lock(mu_);
    const long long t = next_++;
    queue_.push_back({t, Request{sql, std::chrono::steady_clock::now()}});
    cv_.notify_one();
    return t;
}

void QueryServer::loop() {
    for (;;) {
        std::pair<long long, Request> item;
        {
            std::unique_lock<std::mutex> lock(mu_);
            cv_.wait(lock, [&] { return stop_ || !queue_.empty(); });
            if (stop_ && queue_.empty()) return;
            item = std::move(queue_.front());
            queue_.pop_front();
        }
        Done d;
        try {
            d.result = db_.executeQuery(item.second.sql);
        } catch (const std::exception& e) {
            d.rc = -3;
            d.err = e.what();
        }
        d.latency_ms = elapsedMs(item.second.arrival, std::chrono::steady_clock::now());
        {
            std::lock_guard<std::mutex> lock(mu_);
            latencies_.push_back(d.latency_ms);
            done_[item.first] = std::move(d);
        }
        done_cv_.notify_all();
    }
}

int QueryServer::wait(long long ticket, int timeout_ms, QueryResult& out, std::string& err) {
    std::unique_lock<std::mutex> lock(mu_);
    auto ready = [&] { return done_.count(ticket) > 0; };
    if (timeout_ms < 0) done_cv_.wait(lock, ready);
    else if (!done_cv_.wait_for(lock, std::chrono::milliseconds(timeout_ms), ready)) {
        err = "timeout";
        return -3;
    }
    Done d = std::move(done_[ticket]);
    done_.erase(ticket);
    if (d.rc != 0) {
        err = d.err;
        return d.rc;
    }
    out = std::move(d.result);
    return 0;
}

std::string QueryServer::statsJson() {
    std::lock_guard<std::mutex> lock(mu_);
    std::vector<double> l = latencies_;
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) { return l.empty() ? 0.0 : l[std::min(l.size() - 1, static_cast<size_t>(q * l.size()))]; };
    std::ostringstream o;
    o << "{\"completed\":" << l.size() << ",\"p50_ms\":" << pct(0.50) << ",\"p99_ms\":" << pct(0.99) << "}";
    return o.str();
}

}  // namespace db
