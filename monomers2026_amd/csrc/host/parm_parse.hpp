// parm_parse.hpp -- reader for the Model.SELM key/value file.
//
// Semantics of the reference reader (cuda/SELM_ParmParse.cuh:19-101):
// - A '#' starts a comment that runs to the end of the line.
// - A line that is blank after comment removal is skipped.
// - The first whitespace-separated token is the key.  The value is the rest
//   of the line, leading whitespace included, and is parsed with stream
//   extraction.
// - When a key appears twice, the later line wins.
// - get() throws std::runtime_error("Key not found: <key>") for a missing key.
// - query() returns 1 when the key is missing (the value is untouched) and 0
//   when it is found.
// - A fixed-size array is filled element by element.  It throws when the
//   value runs short.
#pragma once

#include <cstddef>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>

namespace selm_host {

class ParmParse {
  public:
    bool load(const std::string &path) {
        std::ifstream in(path);
        if (!in) return false;
        std::string line;
        while (std::getline(in, line)) {
            const size_t hash = line.find('#');
            if (hash != std::string::npos) line.erase(hash);
            std::istringstream ls(line);
            std::string key;
            if (!(ls >> key)) continue;
            std::string rest;
            std::getline(ls, rest);
            kv_[key] = rest;
        }
        return true;
    }

    // Parse `text` as if it were the file's contents (tests, embedded decks).
    void load_string(const std::string &text) {
        std::istringstream in(text);
        std::string line;
        while (std::getline(in, line)) {
            const size_t hash = line.find('#');
            if (hash != std::string::npos) line.erase(hash);
            std::istringstream ls(line);
            std::string key;
            if (!(ls >> key)) continue;
            std::string rest;
            std::getline(ls, rest);
            kv_[key] = rest;
        }
    }

    bool has(const std::string &key) const { return kv_.count(key) != 0; }

    template <typename T>
    void get(const std::string &key, T &out) const {
        auto it = kv_.find(key);
        if (it == kv_.end()) throw std::runtime_error("Key not found: " + key);
        extract(it->second, out, key);
    }

    template <typename T>
    int query(const std::string &key, T &out) const {
        auto it = kv_.find(key);
        if (it == kv_.end()) return 1;
        extract(it->second, out, key);
        return 0;
    }

  private:
    template <typename T>
    static void extract(const std::string &text, T &out, const std::string &key) {
        std::istringstream vs(text);
        if constexpr (std::is_array_v<T>) {
            for (size_t i = 0; i < std::extent_v<T>; i++)
                if (!(vs >> out[i])) throw std::runtime_error("Not enough data to fill the array for key: " + key);
        } else {
            vs >> out;
        }
    }

    std::unordered_map<std::string, std::string> kv_;
};

}  // namespace selm_host
