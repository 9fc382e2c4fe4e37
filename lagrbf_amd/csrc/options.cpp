#include "options.hpp"

#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace lrb {

std::string normalize_option_name(const std::string& name) {
  size_t b = 0;
  while (b < name.size() && (name[b] == '-' || std::isspace((unsigned char)name[b]))) b++;
  size_t e = name.size();
  while (e > b && std::isspace((unsigned char)name[e - 1])) e--;
  return name.substr(b, e - b);
}

void Options::set(const std::string& name, const std::string& value) {
  std::lock_guard<std::mutex> l(mu_);
  db_[normalize_option_name(name)] = value;
}

void Options::clear() {
  std::lock_guard<std::mutex> l(mu_);
  db_.clear();
}

static bool is_option_token(const std::string& t) {
  // "-name" but not a negative number such as "-1.5" or "-.5"
  if (t.size() < 2 || t[0] != '-') return false;
  char c = t[1];
  return !(std::isdigit((unsigned char)c) || c == '.');
}

int Options::insert_string(const std::string& text) {
  std::istringstream in(text);
  std::vector<std::string> tok;
  std::string t;
  while (in >> t) tok.push_back(t);
  std::lock_guard<std::mutex> l(mu_);
  for (size_t i = 0; i < tok.size(); i++) {
    if (!is_option_token(tok[i])) return 62;
    std::string name = normalize_option_name(tok[i]);
    std::string val;
    if (i + 1 < tok.size() && !is_option_token(tok[i + 1])) val = tok[++i];
    db_[name] = val;
  }
  return 0;
}

int Options::insert_file(const std::string& path) {
  std::ifstream f(path);
  if (!f) return 65;  // PETSC_ERR_FILE_OPEN
  std::string line, all;
  while (std::getline(f, line)) {
    size_t h = line.find('#');
    if (h != std::string::npos) line = line.substr(0, h);
    all += line + " ";
  }
  return insert_string(all);
}

std::map<std::string, std::string> Options::snapshot() const {
  std::lock_guard<std::mutex> l(mu_);
  return db_;
}

Options& global_options() {
  static Options o;
  return o;
}

bool OptView::get_string(const std::string& k, std::string& v) const {
  auto it = db.find(k);
  if (it == db.end()) return false;
  v = it->second;
  return true;
}

bool OptView::get_int(const std::string& k, int& v, bool& bad) const {
  std::string s;
  if (!get_string(k, s)) return false;
  char* e = nullptr;
  errno = 0;
  long r = std::strtol(s.c_str(), &e, 10);
  if (s.empty() || *e || errno) { bad = true; return false; }
  v = (int)r;
  return true;
}

bool OptView::get_real(const std::string& k, double& v, bool& bad) const {
  std::string s;
  if (!get_string(k, s)) return false;
  char* e = nullptr;
  double r = std::strtod(s.c_str(), &e);
  if (s.empty() || *e) { bad = true; return false; }
  v = r;
  return true;
}

bool OptView::get_bool(const std::string& k, bool& v, bool& bad) const {
  std::string s;
  if (!get_string(k, s)) return false;
  for (auto& c : s) c = (char)std::tolower((unsigned char)c);
  if (s.empty() || s == "1" || s == "true" || s == "yes" || s == "on") v = true;
  else if (s == "0" || s == "false" || s == "no" || s == "off") v = false;
  else { bad = true; return false; }
  return true;
}

template <typename T, typename P>
static bool parse_list(const std::string& s, std::vector<T>& v, P parse) {
  v.clear();
  std::string cur;
  std::string src = s;
  for (auto& c : src) if (c == '[' || c == ']') c = ' ';
  std::istringstream in(src);
  while (std::getline(in, cur, ',')) {
    size_t b = cur.find_first_not_of(" \t"), e = cur.find_last_not_of(" \t");
    if (b == std::string::npos) return false;
    cur = cur.substr(b, e - b + 1);
    T x;
    if (!parse(cur, x)) return false;
    v.push_back(x);
  }
  return !v.empty();
}

bool OptView::get_int_array(const std::string& k, std::vector<int>& v, bool& bad) const {
  std::string s;
  if (!get_string(k, s)) return false;
  bool ok = parse_list<int>(s, v, [](const std::string& t, int& x) {
    char* e = nullptr; long r = std::strtol(t.c_str(), &e, 10); x = (int)r; return !*e;
  });
  if (!ok) bad = true;
  return ok;
}

bool OptView::get_real_array(const std::string& k, std::vector<double>& v, bool& bad) const {
  std::string s;
  if (!get_string(k, s)) return false;
  bool ok = parse_list<double>(s, v, [](const std::string& t, double& x) {
    char* e = nullptr; x = std::strtod(t.c_str(), &e); return !*e;
  });
  if (!ok) bad = true;
  return ok;
}

}  // namespace lrb
