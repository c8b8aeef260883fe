// skelcl::detail::Event (detail/Event.h, src/Event.cpp:89-98): completion of
// enqueued copies, waited for by host access.  Each entry is a HIP event
// recorded on the library stream after the operation it marks.
#pragma once

#include <memory>
#include <vector>

#include "Runtime.h"

namespace skelcl {
namespace detail {

class Event {
public:
  Event() = default;

  // records a marker after everything enqueued on the library stream so far
  static Event recordNow()
  {
    skl_event e = nullptr;
    check(skl_event_create(&e), "event");
    std::shared_ptr<skl_event_s> p(e, [](skl_event x) { skl_event_destroy(x); });
    check(skl_event_record(e, nullptr), "event record");
    Event ev;
    ev._events.push_back(p);
    return ev;
  }
  void insert(const Event& other)
  {
    _events.insert(_events.end(), other._events.begin(), other._events.end());
  }
  // Event::wait (src/Event.cpp:89-98): blocks until every marker completed
  void wait() const
  {
    for (auto& e : _events) check(skl_event_synchronize(e.get()), "event wait");
  }
  size_t size() const { return _events.size(); }

private:
  std::vector<std::shared_ptr<skl_event_s>> _events;
};

}  // namespace detail
}  // namespace skelcl
