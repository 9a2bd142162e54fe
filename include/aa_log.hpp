// Logging for the dedispersion library: same log_level names and LOG(level,
// msg) macro as the reference (include/aa_log.hpp:30-56), written to stderr
// with a timestamp.  The threshold is read once from AA_LOG_LEVEL
// (error|notice|debug|dev_debug, default notice).
#ifndef ASTRO_ACCELERATE_AA_LOG_HPP
#define ASTRO_ACCELERATE_AA_LOG_HPP

#include <string>

namespace astroaccelerate {

enum class log_level : int { error = 0, warning = 1, notice = 2, debug = 3, dev_debug = 4 };

void aa_log_message(log_level level, const std::string &msg, const char *file, int line);
bool aa_log_enabled(log_level level);

}  // namespace astroaccelerate

#define LOG(level, msg)                                                                  \
    do {                                                                                 \
        if (::astroaccelerate::aa_log_enabled(level))                                   \
            ::astroaccelerate::aa_log_message(level, (msg), __FILE__, __LINE__);        \
    } while (0)

#endif
