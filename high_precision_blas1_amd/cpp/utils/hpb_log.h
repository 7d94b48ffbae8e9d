// hpb_log.h -- minimal logger for the solver templates.
//
// Same role as utils::log_std / utils::logged_obj_base of the reference
// (source/utils/log.h:161-182, logged_obj_base.h:48-128): printf-style info and
// error messages with an object prefix and a verbosity level.
#pragma once
#include <cstdarg>
#include <cstdio>
#include <string>

namespace hpb {

class log_std {
public:
    explicit log_std(int verbosity = 1) : verbosity_(verbosity) {}
    void set_verbosity(int v) { verbosity_ = v; }
    int verbosity() const { return verbosity_; }
    void vmsg(int lev, const std::string& prefix, const char* fmt, va_list ap) const
    {
        if (lev > verbosity_) return;
        std::fprintf(stderr, "%s", prefix.c_str());
        std::vfprintf(stderr, fmt, ap);
        std::fprintf(stderr, "\n");
    }

private:
    int verbosity_;
};

// Base for objects that log with a prefix; a null logger silences them.
template <class Log>
class logged_obj_base {
public:
    logged_obj_base(Log* log = nullptr, int obj_log_lev = 0, const std::string& prefix = "")
        : log_(log), lev_(obj_log_lev), prefix_(prefix)
    {
    }
    void set_log_msg_prefix(const std::string& p) { prefix_ = p; }
    const std::string& get_log_msg_prefix() const { return prefix_; }
    void info_f(const char* fmt, ...) const
    {
        if (!log_) return;
        va_list ap;
        va_start(ap, fmt);
        log_->vmsg(lev_ + 1, prefix_, fmt, ap);
        va_end(ap);
    }
    void error_f(const char* fmt, ...) const
    {
        if (!log_) return;
        va_list ap;
        va_start(ap, fmt);
        log_->vmsg(0, prefix_, fmt, ap);
        va_end(ap);
    }

protected:
    Log* log_;
    int lev_;
    std::string prefix_;
};

}  // namespace hpb
