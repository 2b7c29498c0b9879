// errors.hpp -- last-error message shared by the C ABI translation units.
#pragma once
namespace cdx {
void set_last_error(const char* msg);
}
