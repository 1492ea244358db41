"""dbox_amd: MI355X-native b2World.Step hot path (broadphase, narrowphase, solver) behind a C ABI."""
