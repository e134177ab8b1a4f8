// ConfigFile: the preconditioner config-file API of the reference
// (modules/config_file/config_file.h:54-116, parser config_file.cc:68-142),
// re-implemented with the same names and behaviour:
//   * "key = value" lines; text after '#' is a comment;
//   * a value continues on following lines that hold no delimiter; a blank line,
//     a line with a delimiter or the sentry ends it (continuations join with '\n');
//   * a line containing the sentry "EndConfigFile" stops parsing;
//   * keys and values are trimmed; a repeated key keeps its last value;
//   * read<T>(key) throws key_not_found, read<T>(key, default) returns the default;
//   * the constructor throws file_not_found when the file cannot be opened.
#pragma once

#include <algorithm>
#include <cctype>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>

class ConfigFile {
public:
    struct file_not_found {
        std::string filename;
        explicit file_not_found(const std::string& f = std::string()) : filename(f) {}
    };
    struct key_not_found {
        std::string key;
        explicit key_not_found(const std::string& k = std::string()) : key(k) {}
    };

    ConfigFile(std::string filename, std::string delimiter = "=", std::string comment = "#",
               std::string sentry = "EndConfigFile")
        : delim_(std::move(delimiter)), comment_(std::move(comment)), sentry_(std::move(sentry)) {
        std::ifstream in(filename.c_str());
        if (!in) throw file_not_found(filename);
        parse(in);
    }
    ConfigFile() : delim_("="), comment_("#") {}

    template <class T> T read(const std::string& key) const {
        auto it = kv_.find(key);
        if (it == kv_.end()) throw key_not_found(key);
        return convert<T>(it->second);
    }
    template <class T> T read(const std::string& key, const T& value) const {
        auto it = kv_.find(key);
        return it == kv_.end() ? value : convert<T>(it->second);
    }
    template <class T> bool readInto(T& var, const std::string& key) const {
        auto it = kv_.find(key);
        if (it == kv_.end()) return false;
        var = convert<T>(it->second);
        return true;
    }
    template <class T> bool readInto(T& var, const std::string& key, const T& value) const {
        bool found = readInto(var, key);
        if (!found) var = value;
        return found;
    }
    template <class T> void add(std::string key, const T& value) {
        std::ostringstream os;
        os << value;
        std::string v = os.str();
        kv_[trimmed(key)] = trimmed(v);
    }
    void remove(const std::string& key) { kv_.erase(key); }
    bool keyExists(const std::string& key) const { return kv_.count(key) != 0; }

    std::string getDelimiter() const { return delim_; }
    std::string getComment() const { return comment_; }
    std::string getSentry() const { return sentry_; }
    std::string setDelimiter(const std::string& s) { std::string o = delim_; delim_ = s; return o; }
    std::string setComment(const std::string& s) { std::string o = comment_; comment_ = s; return o; }

    friend std::ostream& operator<<(std::ostream& os, const ConfigFile& cf) {
        for (const auto& kv : cf.kv_) os << kv.first << " " << cf.delim_ << " " << kv.second << std::endl;
        return os;
    }
    friend std::istream& operator>>(std::istream& is, ConfigFile& cf) {
        cf.parse(is);
        return is;
    }

    const std::map<std::string, std::string>& contents() const { return kv_; }

private:
    std::string delim_, comment_, sentry_;
    std::map<std::string, std::string> kv_;

    static std::string trimmed(std::string s) {
        const char* ws = " \n\t\v\r\f";
        size_t b = s.find_first_not_of(ws);
        if (b == std::string::npos) return std::string();
        size_t e = s.find_last_not_of(ws);
        return s.substr(b, e - b + 1);
    }
    std::string strip_comment(const std::string& s) const { return s.substr(0, s.find(comment_)); }
    bool has_sentry(const std::string& s) const { return !sentry_.empty() && s.find(sentry_) != std::string::npos; }

    template <class T> static T convert(const std::string& s) {
        T t{};
        std::istringstream is(s);
        is >> t;
        return t;
    }

    void parse(std::istream& is) {
        std::string pending;      // a line read ahead while collecting a multi-line value
        bool have_pending = false;
        while (true) {
            std::string line;
            if (have_pending) {
                line = pending;
                have_pending = false;
            } else if (!std::getline(is, line)) {
                break;
            }
            line = strip_comment(line);
            if (has_sentry(line)) return;
            size_t pos = line.find(delim_);
            if (pos == std::string::npos) continue;
            std::string key = line.substr(0, pos);
            std::string value = line.substr(pos + delim_.size());
            // gather continuation lines
            std::string next;
            while (std::getline(is, next)) {
                if (trimmed(next).empty()) { pending = next; have_pending = !next.empty(); break; }
                std::string body = strip_comment(next);
                if (body.find(delim_) != std::string::npos || has_sentry(body)) {
                    pending = body;
                    have_pending = !body.empty();
                    break;
                }
                if (!trimmed(body).empty()) value += "\n";
                value += body;
            }
            kv_[trimmed(key)] = trimmed(value);
        }
    }
};

template <> inline std::string ConfigFile::convert<std::string>(const std::string& s) { return s; }
template <> inline bool ConfigFile::convert<bool>(const std::string& s) {
    std::string up = s;
    std::transform(up.begin(), up.end(), up.begin(), [](unsigned char c) { return (char)std::toupper(c); });
    return !(up == "FALSE" || up == "F" || up == "NO" || up == "N" || up == "0" || up == "NONE");
}
