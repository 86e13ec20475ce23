// Minimal JSON reader for Marvin network files (core/3dmatch-net-test.json).
// Recursive descent over the full JSON grammar; numbers are kept as double.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace tdm {

struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;  // keeps file order

    const Json *get(const std::string &key) const {
        for (auto &kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

// Parses `text`; on failure returns false and a message with the offset.
bool json_parse(const std::string &text, Json &out, std::string &err);

}  // namespace tdm
