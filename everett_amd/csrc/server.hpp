// Continuous-batching server entry points (server.cpp).
#pragma once

#include "circuit.hpp"
#include "common.hpp"

namespace ev {

ev_server *server_create(const ev_server_config &cfg);
void server_destroy(ev_server *s);
uint64_t server_submit(ev_server *s, const Circuit &c, uint64_t shots, uint64_t seed);
bool server_poll(ev_server *s, uint64_t id);
ev_request_info server_wait(ev_server *s, uint64_t id, uint64_t *samples);
ev_server_stats server_stats(ev_server *s);

}  // namespace ev
