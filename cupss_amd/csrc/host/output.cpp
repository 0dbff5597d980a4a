// output.cpp -- background writer of field output files (see output.hpp).
#include "output.hpp"

#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <iostream>
#include <mutex>
#include <thread>

namespace cupss_host {

namespace {

// The reference's writeToFile loop: header "x, [y, [z, ]]name", then one
// line per point "%i, [%i, [%i, ]]%.<precision>f" with x fastest.
void write_file(const OutputJob &j)
{
    FILE *fp = std::fopen(j.file.c_str(), "w+");
    if (fp == nullptr) {
        std::cout << "Error creating output file " << j.file << std::endl;
        std::_Exit(1);
    }
    std::string fmt = "%i, ";
    std::fprintf(fp, "x, ");
    if (j.dim == 2) {
        std::fprintf(fp, "y, ");
        fmt += "%i, ";
    }
    if (j.dim == 3) {
        std::fprintf(fp, "y, z, ");
        fmt += "%i, %i, ";
    }
    std::fprintf(fp, "%s\n", j.name.c_str());
    fmt += "%." + std::to_string(j.precision) + "f\n";
    for (int k = 0; k < j.sz; k++)
        for (int jy = 0; jy < j.sy; jy++)
            for (int i = 0; i < j.sx; i++) {
                const float v = j.values[(size_t)k * j.sx * j.sy + (size_t)jy * j.sx + i];
                int w = 0;
                if (j.dim == 1) w = std::fprintf(fp, fmt.c_str(), i, v);
                if (j.dim == 2) w = std::fprintf(fp, fmt.c_str(), i, jy, v);
                if (j.dim == 3) w = std::fprintf(fp, fmt.c_str(), i, jy, k, v);
                if (w < 0) {
                    std::cout << "Error writing data to output file " << j.file << std::endl;
                    std::_Exit(1);
                }
            }
    std::fclose(fp);
}

class Writer {
public:
    static constexpr size_t MAX_PENDING = 8;

    void submit(OutputJob &&job)
    {
        std::unique_lock<std::mutex> lk(m_);
        if (!started_) {
            worker_ = std::thread([this] { loop(); });
            started_ = true;
        }
        space_.wait(lk, [this] { return q_.size() < MAX_PENDING; });
        q_.push_back(std::move(job));
        work_.notify_one();
    }

    void flush()
    {
        std::unique_lock<std::mutex> lk(m_);
        idle_.wait(lk, [this] { return q_.empty() && !busy_; });
    }

    void shutdown()
    {
        {
            std::unique_lock<std::mutex> lk(m_);
            if (!started_) return;
            idle_.wait(lk, [this] { return q_.empty() && !busy_; });
            stop_ = true;
            work_.notify_one();
        }
        worker_.join();
        started_ = false;
    }

private:
    void loop()
    {
        for (;;) {
            OutputJob job;
            {
                std::unique_lock<std::mutex> lk(m_);
                work_.wait(lk, [this] { return stop_ || !q_.empty(); });
                if (q_.empty()) return;  // stop_ with nothing left
                job = std::move(q_.front());
                q_.pop_front();
                busy_ = true;
                space_.notify_one();
            }
            write_file(job);
            {
                std::lock_guard<std::mutex> lk(m_);
                busy_ = false;
                if (q_.empty()) idle_.notify_all();
            }
        }
    }

    std::mutex m_;
    std::condition_variable work_, space_, idle_;
    std::deque<OutputJob> q_;
    std::thread worker_;
    bool started_ = false, stop_ = false, busy_ = false;
};

Writer &writer()
{
    // lives until process exit; drained by an exit handler registered once
    static Writer *w = [] {
        Writer *p = new Writer();
        std::atexit([] { writer().shutdown(); });
        return p;
    }();
    return *w;
}

bool synchronous()
{
    const char *e = std::getenv("CUPSS_SYNC_OUTPUT");
    return e && e[0] == '1';
}

}  // namespace

void output_submit(OutputJob &&job)
{
    if (synchronous()) {
        write_file(job);
        return;
    }
    writer().submit(std::move(job));
}

void output_flush()
{
    if (!synchronous()) writer().flush();
}

}  // namespace cupss_host
