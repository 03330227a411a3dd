// tune.hpp - process-wide cache of autotuned launch choices, persisted as a
// flat JSON object {"key": "value", ...} in $NUMA_HIP_TUNE_CACHE (default
// ~/.cache/numa_hip/tune.json).  Keys name the device, the kernel family and
// the problem shape; values are the chosen parameters.  A cache that cannot
// be read or written is ignored (the choices are then re-timed per process).
#pragma once

#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/file.h>
#include <unistd.h>

#include <atomic>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <set>
#include <sstream>
#include <string>

namespace numa_hip {

class TuneCache {
public:
    static TuneCache& get() {
        static TuneCache c;
        return c;
    }
    bool lookup(const std::string& key, std::string& val) {
        std::lock_guard<std::mutex> g(mu_);
        auto it = m_.find(key);
        if (it == m_.end()) return false;
        val = it->second;
        return true;
    }
    void store(const std::string& key, const std::string& val) {
        std::lock_guard<std::mutex> g(mu_);
        m_[key] = val;
        mine_.insert(key);
        save();
    }
    std::string path() const { return path_; }

    // Held while one key is tuned: solvers of several threads that need the
    // same key wait for the first one's choice instead of timing their own
    // (under each other's load), so every solver of the process runs the
    // same choice for a key.
    std::unique_lock<std::mutex> tuning(const std::string& key) {
        std::mutex* m = nullptr;
        {
            std::lock_guard<std::mutex> g(mu_);
            auto& slot = key_mu_[key];
            if (!slot) slot = std::make_unique<std::mutex>();
            m = slot.get();
        }
        return std::unique_lock<std::mutex>(*m);
    }

    // "<arch>|<CUs>cu" of a device
    static std::string device_key(int dev) {
        hipDeviceProp_t p{};
        if (hipGetDeviceProperties(&p, dev) != hipSuccess) return "unknown";
        std::ostringstream o;
        o << p.gcnArchName << "|" << p.multiProcessorCount << "cu";
        return o.str();
    }

private:
    TuneCache() {
        if (const char* e = std::getenv("NUMA_HIP_TUNE_CACHE"); e && *e) {
            path_ = e;
        } else if (const char* h = std::getenv("HOME"); h && *h) {
            path_ = std::string(h) + "/.cache/numa_hip/tune.json";
        }
        load();
    }
    void load() { parse(path_, m_); }
    // strings between quotes, alternating key / value
    static void parse(const std::string& path, std::map<std::string, std::string>& out) {
        if (path.empty()) return;
        std::ifstream f(path);
        if (!f) return;
        std::stringstream ss;
        ss << f.rdbuf();
        const std::string s = ss.str();
        std::string cur, key;
        bool in = false, have_key = false;
        for (char ch : s) {
            if (ch == '"') {
                if (in) {
                    if (!have_key) {
                        key = cur;
                        have_key = true;
                    } else {
                        out[key] = cur;
                        have_key = false;
                    }
                    cur.clear();
                }
                in = !in;
            } else if (in) {
                cur.push_back(ch);
            }
        }
    }
    // Processes that tune at once (a benchmark beside a test run) share the
    // file: under an exclusive flock on "<path>.lock" the file is re-read, the
    // entries of other processes merged in (the keys this process tuned itself keep its values),
    // and the result written to a temporary name unique to this process and
    // renamed over the file, so no reader ever sees a partial file.
    void save() {
        if (path_.empty()) return;
        std::error_code ec;
        const std::filesystem::path p(path_);
        if (p.has_parent_path()) std::filesystem::create_directories(p.parent_path(), ec);
        const int fd = ::open((path_ + ".lock").c_str(), O_CREAT | O_RDWR | O_CLOEXEC, 0644);
        if (fd >= 0) (void)::flock(fd, LOCK_EX);
        std::map<std::string, std::string> disk;
        parse(path_, disk);
        for (const auto& [k, v] : disk)
            if (!mine_.count(k)) m_[k] = v;
        static std::atomic<unsigned long> seq{0};
        const std::string tmp = path_ + ".tmp." + std::to_string((long)::getpid()) + "." +
                                std::to_string(seq.fetch_add(1)) + "." +
                                std::to_string(std::random_device{}());
        bool ok = false;
        {
            std::ofstream f(tmp, std::ios::trunc);
            if (f) {
                f << "{\n";
                size_t i = 0;
                for (const auto& [k, v] : m_)
                    f << "  \"" << k << "\": \"" << v << "\"" << (++i < m_.size() ? ",\n" : "\n");
                f << "}\n";
                f.flush();
                ok = f.good();
            }
        }
        if (ok) std::filesystem::rename(tmp, path_, ec);
        if (!ok || ec) std::filesystem::remove(tmp, ec);
        if (fd >= 0) {
            (void)::flock(fd, LOCK_UN);
            (void)::close(fd);
        }
    }

    std::mutex mu_;
    std::map<std::string, std::string> m_;
    std::set<std::string> mine_;  // keys this process has tuned itself
    std::map<std::string, std::unique_ptr<std::mutex>> key_mu_;
    std::string path_;
};

}  // namespace numa_hip
