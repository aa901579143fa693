{
  "targets": [
    {
      "target_name": "bd",
      "sources": ["bd_node.c"],
      "include_dirs": ["<(module_root_dir)/../../include"],
      "cflags": ["-std=c11", "-Wall"],
      "libraries": ["-L<(module_root_dir)/..", "-lbd", "-Wl,-rpath,'$$ORIGIN/../../..'"]
    }
  ]
}
