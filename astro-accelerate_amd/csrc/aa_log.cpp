// stderr logger behind the LOG macro.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>

#include "aa_log.hpp"

namespace astroaccelerate {

namespace {
log_level threshold() {
    static const log_level level = [] {
        const char *env = std::getenv("AA_LOG_LEVEL");
        if (!env) return log_level::notice;
        if (!std::strcmp(env, "error")) return log_level::error;
        if (!std::strcmp(env, "warning")) return log_level::warning;
        if (!std::strcmp(env, "debug")) return log_level::debug;
        if (!std::strcmp(env, "dev_debug")) return log_level::dev_debug;
        return log_level::notice;
    }();
    return level;
}
const char *level_name(log_level l) {
    switch (l) {
        case log_level::error: return "ERROR";
        case log_level::warning: return "WARNING";
        case log_level::notice: return "NOTICE";
        case log_level::debug: return "DEBUG";
        case log_level::dev_debug: return "DEV_DEBUG";
    }
    return "?";
}
}  // namespace

bool aa_log_enabled(log_level level) { return (int)level <= (int)threshold(); }

void aa_log_message(log_level level, const std::string &msg, const char *file, int line) {
    const std::time_t now = std::time(nullptr);
    char stamp[32];
    std::strftime(stamp, sizeof stamp, "%Y-%m-%d %H:%M:%S", std::localtime(&now));
    const char *base = std::strrchr(file, '/');
    std::fprintf(stderr, "[%s] %s %s:%d %s\n", stamp, level_name(level), base ? base + 1 : file, line,
                 msg.c_str());
}

}  // namespace astroaccelerate
