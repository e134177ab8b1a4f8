// Error plumbing shared by the host construction and the C ABI.
#pragma once

#include <stdexcept>
#include <string>

namespace mlp {

// A failure with the status code the C ABI returns for it.
class Error : public std::runtime_error {
public:
    explicit Error(const std::string& msg, int code = 6) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }
private:
    int code_;
};

void set_last_error(const std::string& msg);

}  // namespace mlp

#define MLP_HIP_CHECK(expr)                                                                  \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw ::mlp::Error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " + \
                                   __FILE__ + ":" + std::to_string(__LINE__) + " (" #expr ")", \
                               2);                                                           \
    } while (0)
