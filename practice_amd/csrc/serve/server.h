// Query server: request queue in front of the database.
#pragma once

#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>

#include "../runtime/database.h"

namespace db {

class QueryServer {
public:
    QueryServer(DatabaseInstance& db, const std::string& options);
    ~QueryServer();
    long long submit(const std::string& sql);
    // 0 on success, negative error code otherwise (err receives the message)
    int wait(long long ticket, int timeout_ms, QueryResult& out, std::string& err);
    std::string statsJson();

private:
    struct Request {
        std::string sql;
        std::chrono::steady_clock::time_point arrival;
    };
    struct Done {
        int rc = 0;
        std::string err;
        QueryResult result;
        double latency_ms = 0;
    };
    DatabaseInstance& db_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::deque<std::pair<long long, Request>> queue_;
    std::map<long long, Done> done_;
    long long next_ = 0;
    bool stop_ = false;
    std::thread worker_;
    std::vector<double> latencies_;
    void loop();
};

}  // namespace db
