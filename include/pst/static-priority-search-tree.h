// Abstract base of the tree classes, with the reference's interface
// (pst/static-priority-search-tree.h:9-28): a pure-virtual print and an
// operator<< that calls it. Searches are member templates of the derived
// classes, so they are not virtual.
#ifndef PST_AMD_STATIC_PRIORITY_SEARCH_TREE_H
#define PST_AMD_STATIC_PRIORITY_SEARCH_TREE_H

#include <cstddef>
#include <ostream>

template <typename T, template <typename, typename, size_t> class PointStructTemplate, typename IDType = void,
          size_t num_IDs = 0>
class StaticPrioritySearchTree {
 public:
  virtual ~StaticPrioritySearchTree() = default;
  virtual void print(std::ostream& os) const = 0;
};

template <typename T, template <typename, typename, size_t> class PointStructTemplate,
          template <typename, template <typename, typename, size_t> class, typename, size_t> class Tree,
          typename IDType, size_t num_IDs>
std::ostream& operator<<(std::ostream& os, const Tree<T, PointStructTemplate, IDType, num_IDs>& t) {
  t.print(os);
  return os;
}

#endif
