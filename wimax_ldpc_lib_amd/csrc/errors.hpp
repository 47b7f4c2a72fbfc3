// errors.hpp -- the calling thread's last-error text behind wl_last_error().
#pragma once

namespace wl {
void set_error_text(const char* text);
}
