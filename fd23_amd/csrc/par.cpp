// Par-file / command-line parameters with the semantics of the Caltech
// getpar the reference drivers use (PSV/clean_nobox/getpar/getpar.c):
//   - argv entries "name=value"; "-name" records the switch "SWITCH=name";
//     "par=file" reads that file at that point (nested up to 8 levels);
//     "ENDPAR" stops argv processing (setpar, :145-290)
//   - par files hold any number of name=value entries per line separated by
//     blanks or tabs, '#' starts a comment, values may be quoted with ' or "
//     (\" inside), a backslash before the newline continues the line
//     (gp_do_par_file :632-722, gp_fgets :877-912)
//   - "$(name)" inside a value is replaced by the value of an earlier entry
//     (gp_subpar :800-875)
//   - the last occurrence of a name wins (backward search, :490-491)
#include "fd23.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

extern "C" void fd23_set_error_(const char *msg);

struct fd23_par {
    std::vector<std::pair<std::string, std::string>> list;
    std::string err;

    const std::string *find(const std::string &name) const {
        for (size_t i = list.size(); i-- > 0;)
            if (list[i].first == name) return &list[i].second;
        return nullptr;
    }
};

namespace {

// expand "$(name)" at s[i] (s[i] == '$'); returns the index after it
size_t substitute(const fd23_par &p, const std::string &s, size_t i, std::string &out) {
    if (i > 0 && s[i - 1] == '\\') {
        out.back() = '$';
        return i + 1;
    }
    if (i + 1 < s.size() && s[i + 1] == '(') {
        std::string name;
        size_t j = i + 2;
        while (j < s.size() && s[j] != ')') {
            if (s[j] == '$') j = substitute(p, s, j, name);
            else name += s[j++];
        }
        if (j < s.size() && s[j] == ')') {
            if (const std::string *v = p.find(name)) out += *v;
            return j + 1;
        }
    }
    out += '$';
    return i + 1;
}

// parse a value starting at s[i]; stop set depends on the source
std::string read_value(const fd23_par &p, const std::string &s, size_t &i, bool from_file) {
    std::string v;
    if (i < s.size() && (s[i] == '"' || s[i] == '\'')) {
        const char t = s[i++];
        while (i < s.size() && s[i] != '\n') {
            if (s[i] == t) {
                if (i == 0 || s[i - 1] != '\\') {
                    i++;
                    break;
                }
                v.back() = t;
                i++;
            } else if (s[i] == '$') {
                i = substitute(p, s, i, v);
            } else {
                v += s[i++];
            }
        }
        return v;
    }
    while (i < s.size()) {
        const char ch = s[i];
        if (from_file && (ch == '\n' || ch == '\t' || ch == ' ')) break;
        if (ch == '$') i = substitute(p, s, i, v);
        else v += s[i++];
    }
    return v;
}

int do_file(fd23_par &p, const std::string &fname, int level);

void add(fd23_par &p, const std::string &name, const std::string &value, int level, int &rc) {
    p.list.emplace_back(name, value);
    if (name == "par" && rc == 0) rc = do_file(p, value, level + 1);
}

int do_file(fd23_par &p, const std::string &fname, int level) {
    if (level > 8) {
        p.err = "too many recursive par files";
        return FD23_EINVAL;
    }
    if (fname.empty()) return FD23_OK;
    FILE *f = fopen(fname.c_str(), "r");
    if (!f) {
        p.err = "cannot open par file " + fname;
        return FD23_EIO;
    }
    std::string text;
    char buf[4096];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
    fclose(f);
    // logical lines: backslash-newline joins
    std::vector<std::string> lines;
    std::string cur;
    for (size_t i = 0; i < text.size(); i++) {
        if (text[i] == '\n') {
            if (!cur.empty() && cur.back() == '\\') {
                cur.pop_back();
                continue;
            }
            lines.push_back(cur);
            cur.clear();
        } else if (text[i] == '\0') {
            break;
        } else {
            cur += text[i];
        }
    }
    if (!cur.empty()) lines.push_back(cur);
    int rc = FD23_OK;
    for (const std::string &line : lines) {
        size_t i = 0;
        while (rc == FD23_OK) {
            while (i < line.size() && (line[i] == ' ' || line[i] == '\t')) i++;
            if (i >= line.size() || line[i] == '#') break;
            std::string name;
            while (i < line.size() && line[i] != '=' && line[i] != ' ' && line[i] != '\t') name += line[i++];
            if (i < line.size() && line[i] == '=') i++;
            std::string value = read_value(p, line, i, true);
            add(p, name, value, level, rc);
        }
        if (rc) break;
    }
    return rc;
}

} // namespace

extern "C" int fd23_par_open(int argc, const char *const *argv, fd23_par **out) {
    if (!out) return FD23_EINVAL;
    fd23_par *p = new fd23_par();
    int rc = FD23_OK;
    for (int a = 1; a < argc && rc == FD23_OK; a++) {
        std::string s = argv[a] ? argv[a] : "";
        size_t i = 0;
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) i++;
        std::string name;
        while (i < s.size() && s[i] != '=') name += s[i++];
        if (i < s.size() && s[i] == '=') i++;
        std::string value = read_value(*p, s, i, false);
        if (!name.empty() && name[0] == '-') p->list.emplace_back("SWITCH", name.substr(1));
        else add(*p, name, value, 0, rc);
        if (name == "ENDPAR") break;
    }
    if (rc) {
        fd23_set_error_(p->err.c_str());
        delete p;
        *out = nullptr;
        return rc;
    }
    *out = p;
    return FD23_OK;
}

extern "C" void fd23_par_close(fd23_par *p) { delete p; }

extern "C" int fd23_par_int(fd23_par *p, const char *name, int *val) {
    if (!p || !name) return FD23_EINVAL;
    const std::string *v = p->find(name);
    if (!v) return 0;
    if (val) *val = atoi(v->c_str());
    return 1;
}
extern "C" int fd23_par_float(fd23_par *p, const char *name, float *val) {
    if (!p || !name) return FD23_EINVAL;
    const std::string *v = p->find(name);
    if (!v) return 0;
    if (val) *val = (float)atof(v->c_str());
    return 1;
}
extern "C" int fd23_par_double(fd23_par *p, const char *name, double *val) {
    if (!p || !name) return FD23_EINVAL;
    const std::string *v = p->find(name);
    if (!v) return 0;
    if (val) *val = atof(v->c_str());
    return 1;
}
extern "C" int fd23_par_string(fd23_par *p, const char *name, char *buf, int cap) {
    if (!p || !name) return FD23_EINVAL;
    const std::string *v = p->find(name);
    if (!v) return 0;
    if (buf && cap > 0) {
        if ((int)v->size() >= cap) return FD23_EINVAL;
        memcpy(buf, v->c_str(), v->size() + 1);
    }
    return 1;
}
