// Program registry: name -> kernel host stub.
#include "registry.hpp"

#include <cstring>
#include <vector>

namespace gpumpi {
namespace progs {

namespace {
struct Entry {
    const char* name;
    const void* kernel;
};
std::vector<Entry>& table() {
    static std::vector<Entry> t;
    return t;
}
}  // namespace

Registrar::Registrar(const char* name, const void* kernel) { table().push_back({name, kernel}); }

}  // namespace progs
}  // namespace gpumpi

extern "C" {
// Kernel of the program called `name`, or NULL.
const void* gpumpi_prog_lookup(const char* name) {
    for (const auto& e : gpumpi::progs::table())
        if (std::strcmp(e.name, name) == 0) return e.kernel;
    return nullptr;
}
int gpumpi_prog_count(void) { return (int)gpumpi::progs::table().size(); }
const char* gpumpi_prog_name(int i) {
    if (i < 0 || i >= (int)gpumpi::progs::table().size()) return nullptr;
    return gpumpi::progs::table()[(size_t)i].name;
}
}
