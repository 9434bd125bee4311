// Host check of guard.h (the C entry points' exception barrier): a body that throws
// std::bad_alloc returns SHG_E_NOMEM, any other exception SHG_E_INTERNAL, and a normal
// body its own return code; the message names the entry point.  Prints one line per case.
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../shg_focused_gaussian_beams_amd/csrc/guard.h"

static std::string g_msg;
int shg_internal_fail(int code, const std::string &msg) {
    g_msg = msg;
    return code;
}

int main() {
    const int a = shg_internal_guard("entry_a", []() -> int { throw std::bad_alloc(); });
    std::printf("%d %s\n", a, g_msg.c_str());
    const int b = shg_internal_guard("entry_b", []() -> int { throw std::length_error("vector"); });
    std::printf("%d %s\n", b, g_msg.c_str());
    const int c = shg_internal_guard("entry_c", []() -> int { throw 7; });
    std::printf("%d %s\n", c, g_msg.c_str());
    const int d = shg_internal_guard("entry_d", []() -> int { return SHG_E_STATE; });
    std::printf("%d\n", d);
    return 0;
}
