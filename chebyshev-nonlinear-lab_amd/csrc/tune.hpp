// tune.hpp - process-wide cache of autotuned launch choices, persisted as a
// flat JSON object {"key": "value", ...} in $NUMA_HIP_TUNE_CACHE (default
// ~/.cache/numa_hip/tune.json).  Keys name the device, the kernel family and
// the problem shape; values are the chosen parameters.  A cache that cannot
// be read or written is ignored (the choices are then re-timed per process).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <map>
#include <mutex>
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
        save();
    }
    std::string path() const { return path_; }

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
    void load() {
        if (path_.empty()) return;
        std::ifstream f(path_);
        if (!f) return;
        std::stringstream ss;
        ss << f.rdbuf();
        const std::string s = ss.str();
        // strings between quotes, alternating key / value
        std::string cur, key;
        bool in = false, have_key = false;
        for (char ch : s) {
            if (ch == '"') {
                if (in) {
                    if (!have_key) {
                        key = cur;
                        have_key = true;
                    } else {
                        m_[key] = cur;
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
    void save() {
        if (path_.empty()) return;
        std::error_code ec;
        const std::filesystem::path p(path_);
        if (p.has_parent_path()) std::filesystem::create_directories(p.parent_path(), ec);
        const std::string tmp = path_ + ".tmp";
        {
            std::ofstream f(tmp, std::ios::trunc);
            if (!f) return;
            f << "{\n";
            size_t i = 0;
            for (const auto& [k, v] : m_) f << "  \"" << k << "\": \"" << v << "\"" << (++i < m_.size() ? ",\n" : "\n");
            f << "}\n";
        }
        std::filesystem::rename(tmp, path_, ec);
    }

    std::mutex mu_;
    std::map<std::string, std::string> m_;
    std::string path_;
};

}  // namespace numa_hip
