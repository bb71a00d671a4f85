// Host-side error messages that carry their C-ABI status.
//
// A failing host call returns a std::string; a message made with tagged() starts with a two-byte tag
// (0x01 and the status digit) that the C ABI turns into the csm_status it returns and strips from
// the text csm_last_error() shows.  An untagged message is a HIP runtime failure (CSM_ERR_DEVICE).
#pragma once
#include <string>

namespace csm {

enum class Code : char {
    kArgument = '1',      // CSM_ERR_ARGUMENT
    kIo = '2',            // CSM_ERR_IO
    kFormat = '3',        // CSM_ERR_FORMAT
    kLimit = '4',         // CSM_ERR_LIMIT
    kDevice = '5',        // CSM_ERR_DEVICE
    kMemory = '6',        // CSM_ERR_MEMORY
    kState = '7',         // CSM_ERR_STATE
    kPrecondition = '8',  // CSM_ERR_PRECONDITION
    kMismatch = '9',      // CSM_ERR_MISMATCH
};

inline std::string tagged(Code c, const std::string& msg) {
    std::string s;
    s.reserve(msg.size() + 2);
    s += '\x01';
    s += static_cast<char>(c);
    s += msg;
    return s;
}

// The status digit of a message ('5' for an untagged one) and its text without the tag.
inline char code_of(const std::string& msg) { return msg.size() >= 2 && msg[0] == '\x01' ? msg[1] : static_cast<char>(Code::kDevice); }
inline std::string text_of(const std::string& msg) { return msg.size() >= 2 && msg[0] == '\x01' ? msg.substr(2) : msg; }

}  // namespace csm
