extern "C" int mi_rb_mlp_placeholder() { return 0; }
