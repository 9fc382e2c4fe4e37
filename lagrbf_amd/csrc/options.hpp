// PETSc-free options database: the same option names the reference reads through
// PetscOptionsGet* (lagrangian_rbf.c:135-195, partition_of_unity.c:68-111,
// dormand_prince.c:340-353), filled by lagrbf_set_option(), an options file or the
// LAGRBF_OPTIONS environment variable.
#pragma once
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace lrb {

class Options {
 public:
  void set(const std::string& name, const std::string& value);
  void clear();
  int insert_string(const std::string& text);   // "-a 1 -b x,y -flag"
  int insert_file(const std::string& path);     // PETSc options-file syntax, '#' comments
  std::map<std::string, std::string> snapshot() const;

 private:
  mutable std::mutex mu_;
  std::map<std::string, std::string> db_;
};

Options& global_options();

// Typed lookups on a snapshot.  Each returns true when the option is present and
// parses; `bad` is set when present but malformed.
struct OptView {
  std::map<std::string, std::string> db;
  bool has(const std::string& k) const { return db.count(k) != 0; }
  bool get_int(const std::string& k, int& v, bool& bad) const;
  bool get_real(const std::string& k, double& v, bool& bad) const;
  bool get_bool(const std::string& k, bool& v, bool& bad) const;
  bool get_string(const std::string& k, std::string& v) const;
  bool get_int_array(const std::string& k, std::vector<int>& v, bool& bad) const;
  bool get_real_array(const std::string& k, std::vector<double>& v, bool& bad) const;
};

std::string normalize_option_name(const std::string& name);

}  // namespace lrb
